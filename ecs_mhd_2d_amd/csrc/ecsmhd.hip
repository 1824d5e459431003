// Host orchestration and C-ABI of the fused MHD kernels (gfx950, FP64).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>
#include <stdlib.h>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include "rhs_kernels.h"
#include "adjoint_kernels.h"
#include "vec_kernels.h"
#include "../../include/ecsmhd.h"

using namespace ecs;

static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
#define HIPCHK(x)                                                                           \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return fail(2, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define LAUNCHCHK() HIPCHK(hipGetLastError())

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
#define GUARD(c)                                                                              \
  CtxGuard guard_(&(c)->busy);                                                                \
  if (!guard_.ok) return fail(4, "context is in use by another call (use one context per thread)")

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
  double2* xtape = nullptr;    // [ckpt][B][2][R][KC]
  double2* vtape = nullptr;    // [ckpt][4][B][2][R][KC]
  size_t vt_cap = 0;           // elements of xtape + vtape
  double2* itape = nullptr;    // primal column outputs [ckpt][4][B][8][N][KC]
  size_t it_cap = 0;
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
  size_t compact_elems() const {
    int kmax = (n - 1) / 3;
    return (size_t)2 * (2 * kmax + 1) * (kmax + 1);
  }
};

// Allocation, frees and synchronous copies must not run while another thread
// captures a graph (they would join the capture's dependencies), so both take
// this lock.
static std::recursive_mutex g_alloc_mu;

static void drop_graphs(ecs_ctx* c) {
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
}

static int ensure_cap(ecs_ctx* c, int B) {
  if (B <= c->cap) return 0;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  drop_graphs(c);
  int nc = std::max(B, c->cap * 2);
  const int kmax = (c->n - 1) / 3, KC = kmax + 1;
  const size_t comp = c->compact_elems();
  const size_t inter = (size_t)16 * c->n * KC;
  const size_t rowo = (size_t)4 * c->n * KC;
  double2** arrs[] = {&c->x0, &c->acc, &c->v1, &c->v4, &c->v5, &c->xf, &c->out,
                      &c->dx0, &c->dacc, &c->dv1, &c->dv4, &c->dxf, &c->dout};
  for (auto p : arrs) {
    if (*p) HIPCHK(hipFree(*p));
    HIPCHK(hipMalloc((void**)p, comp * nc * sizeof(double2)));
  }
  if (c->I) HIPCHK(hipFree(c->I));
  if (c->O) HIPCHK(hipFree(c->O));
  HIPCHK(hipMalloc((void**)&c->I, inter * nc * sizeof(double2)));
  HIPCHK(hipMalloc((void**)&c->O, rowo * nc * sizeof(double2)));
  if (c->scal) HIPCHK(hipFree(c->scal));
  if (c->hscal) HIPCHK(hipHostFree(c->hscal));
  HIPCHK(hipMalloc((void**)&c->scal, 4 * nc * sizeof(double)));
  HIPCHK(hipHostMalloc((void**)&c->hscal, 4 * nc * sizeof(double), hipHostMallocDefault));
  if (c->red) HIPCHK(hipFree(c->red));
  if (c->hred) HIPCHK(hipHostFree(c->hred));
  HIPCHK(hipMalloc((void**)&c->red, (size_t)64 * nc * sizeof(double)));
  HIPCHK(hipHostMalloc((void**)&c->hred, (size_t)64 * nc * sizeof(double), hipHostMallocDefault));
  c->cap = nc;
  return 0;
}

static Args base_args(ecs_ctx* c) {
  Args a;
  memset(&a, 0, sizeof(a));
  a.tw = c->tw;
  a.fhat = c->fhat;
  a.nu[0] = c->nu;
  a.nu[1] = c->eta;
  a.b0x = c->b0x;
  a.b0y = c->b0y;
  a.rowout = c->O;
  a.inter = c->I;
  a.x0 = c->x0; a.acc = c->acc; a.v1 = c->v1; a.v4 = c->v4; a.v5 = c->v5; a.xf = c->xf; a.out = c->out;
  a.dx0 = c->dx0; a.dacc = c->dacc; a.dv1 = c->dv1; a.dv4 = c->dv4; a.dxf = c->dxf; a.dout = c->dout;
  const int cap = c->cap;
  a.dt = c->scal;
  a.ddt = c->scal + cap;
  a.Tb = c->scal + 2 * cap;
  a.dTb = c->scal + 3 * cap;
  a.hist = c->hist;
  a.masked_e = 1;
  return a;
}

// ------------------------------------------------------------------ launches
// Tile variants: column kernels take TY in {1, 2, 4} columns per block, row
// kernels TX in {1, 2, 4} rows per block.  The autotuner (autotune<N>) picks
// one per (N, kernel family, batch bucket) by timing each on the device.
enum TuneKind : int { K_COL = 0, K_COLT = 1, K_ROW = 2, K_ROWT = 3 };
static std::mutex g_tune_mu;
static std::map<long long, int> g_tune;  // key -> variant index

static int bucket_of(int B) {
  int k = 0;
  while ((1 << k) < B && k < 12) ++k;
  return k;
}
static long long tune_key(int n, int kind, int B) { return ((long long)n * 8 + kind) * 64 + bucket_of(B); }
static int tuned(int n, int kind, int B, int dflt) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(tune_key(n, kind, B));
  return it == g_tune.end() ? dflt : it->second;
}

// default variant indices (column: TY = 1 << v, row: TX = 1 << v)
template <int N> struct DefTile;
template <> struct DefTile<16> { static constexpr int COL = 1, ROW = 1; };
template <> struct DefTile<32> { static constexpr int COL = 2, ROW = 1; };
template <> struct DefTile<64> { static constexpr int COL = 2, ROW = 1; };
template <> struct DefTile<128> { static constexpr int COL = 1, ROW = 0; };
template <> struct DefTile<256> { static constexpr int COL = 1, ROW = 0; };

template <int N, int TY, int MODE, bool TAN>
static int col_v(const Args& a, int B, hipStream_t st) {
  dim3 grid((Geo<N>::KC + TY - 1) / TY, 2, B);
  hipLaunchKernelGGL((col_kernel<N, TY, NT, MODE, TAN>), grid, dim3(NT), 0, st, a);
  LAUNCHCHK();
  return 0;
}
template <int N, int TX, int RMODE, bool TAN>
static int row_v(const Args& a, int B, hipStream_t st) {
  dim3 grid(N / TX, 1, B);
  hipLaunchKernelGGL((row_kernel<N, TX, NT, RMODE, TAN>), grid, dim3(NT), 0, st, a);
  LAUNCHCHK();
  return 0;
}
template <int N, int MODE, bool TAN>
static int col_var(int v, const Args& a, int B, hipStream_t st) {
  return v == 0 ? col_v<N, 1, MODE, TAN>(a, B, st)
                : (v == 1 ? col_v<N, 2, MODE, TAN>(a, B, st) : col_v<N, 4, MODE, TAN>(a, B, st));
}
template <int N, int RMODE, bool TAN>
static int row_var(int v, const Args& a, int B, hipStream_t st) {
  return v == 0 ? row_v<N, 1, RMODE, TAN>(a, B, st)
                : (v == 1 ? row_v<N, 2, RMODE, TAN>(a, B, st) : row_v<N, 4, RMODE, TAN>(a, B, st));
}
template <int N, int MODE, bool TAN>
static int col(const Args& a, int B, hipStream_t st) {
  return col_var<N, MODE, TAN>(tuned(N, TAN ? K_COLT : K_COL, B, DefTile<N>::COL), a, B, st);
}
template <int N, int RMODE, bool TAN>
static int row(const Args& a, int B, hipStream_t st) {
  return row_var<N, RMODE, TAN>(tuned(N, TAN ? K_ROWT : K_ROW, B, DefTile<N>::ROW), a, B, st);
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

#define TRY(x)              \
  do {                      \
    int r_ = (x);           \
    if (r_) return r_;      \
  } while (0)

static int upload_scalars(ecs_ctx* c, int slot, const double* vals, int B, hipStream_t st) {
  double* h = c->hscal + (size_t)slot * c->cap;
  for (int b = 0; b < B; ++b) h[b] = vals[b];
  HIPCHK(hipMemcpyAsync(c->scal + (size_t)slot * c->cap, h, B * sizeof(double), hipMemcpyHostToDevice, st));
  return 0;
}

// ------------------------------------------------------------------ hipGraph chunks
// Captures the launches `record` makes on the context's own stream once per
// key (the bytes of the kernels' arguments plus a tag), then replays the
// instantiated graph; the caller's stream is joined before and after.
// Opt-in (ECSMHD_GRAPH=1): on this ROCm stack, eager launches of the same
// chunk measured faster, so eager is the default.
static bool use_graphs() { return getenv("ECSMHD_GRAPH") != nullptr; }

static std::string graph_key(const char* tag, const Args& a, int B, int K) {
  std::string k(tag);
  k.append((const char*)&a, sizeof(Args));
  k.append((const char*)&B, sizeof(int));
  k.append((const char*)&K, sizeof(int));
  return k;
}

template <class F>
static int graph_run(ecs_ctx* c, const std::string& key, hipStream_t st, F&& record) {
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

static int set_counter(ecs_ctx* c, int v, hipStream_t st) {
  if (!c->cnt) {
    HIPCHK(hipMalloc((void**)&c->cnt, 2 * sizeof(int)));
    HIPCHK(hipHostMalloc((void**)&c->hcnt, 2 * sizeof(int), hipHostMallocDefault));
  }
  HIPCHK(hipStreamSynchronize(st));  // the pinned word may still feed an earlier copy
  c->hcnt[0] = v;
  c->hcnt[1] = v;
  HIPCHK(hipMemcpyAsync(c->cnt, c->hcnt, 2 * sizeof(int), hipMemcpyHostToDevice, st));
  return 0;
}

// ------------------------------------------------------------------ autotuner
// Times every tile variant of the four RHS kernel families (column, tangent
// column, row, tangent row) at this batch size with HIP events on `st` and
// keeps the fastest for the batch bucket (powers of two).  Runs at the start
// of an entry point, before any input is packed, on zeroed workspace.
// ECSMHD_NO_TUNE=1 keeps the default tiles.
template <int N>
static int autotune(ecs_ctx* c, int B, hipStream_t st) {
  if (getenv("ECSMHD_NO_TUNE")) return 0;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_tune.count(tune_key(N, K_ROWT, B))) return 0;
  }
  const size_t ce = c->compact_elems() * B * sizeof(double2);
  std::vector<double> dt(B, 1e-3);
  TRY(upload_scalars(c, 0, dt.data(), B, st));
  TRY(upload_scalars(c, 1, dt.data(), B, st));
  Args a = base_args(c);
  HIPCHK(hipMemsetAsync(c->x0, 0, ce, st));
  HIPCHK(hipMemsetAsync(c->dx0, 0, ce, st));
  HIPCHK(hipMemsetAsync(c->acc, 0, ce, st));
  HIPCHK(hipMemsetAsync(c->dacc, 0, ce, st));
  TRY((col_var<N, C_LOAD, true>(1, a, B, st)));
  TRY((row_var<N, R_RHS, true>(1, a, B, st)));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int reps = std::max(3, 64 / B);
  float best[4] = {1e30f, 1e30f, 1e30f, 1e30f};
  int pick[4] = {DefTile<N>::COL, DefTile<N>::COL, DefTile<N>::ROW, DefTile<N>::ROW};
  for (int kind = 0; kind < 4; ++kind) {
    const int nv = 3;
    for (int v = 0; v < nv; ++v) {
      auto launch = [&]() -> int {
        switch (kind) {
          case K_COL: return col_var<N, C_S2, false>(v, a, B, st);
          case K_COLT: return col_var<N, C_S2, true>(v, a, B, st);
          case K_ROW: return row_var<N, R_RHS, false>(v, a, B, st);
          default: return row_var<N, R_RHS, true>(v, a, B, st);
        }
      };
      TRY(launch());
      HIPCHK(hipEventRecord(e0, st));
      for (int i = 0; i < reps; ++i) TRY(launch());
      HIPCHK(hipEventRecord(e1, st));
      HIPCHK(hipEventSynchronize(e1));
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, e0, e1));
      if (t < best[kind]) { best[kind] = t; pick[kind] = v; }
    }
  }
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int kind = 0; kind < 4; ++kind) g_tune[tune_key(N, kind, B)] = pick[kind];
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
static int ensure_hist(ecs_ctx* c, size_t need) {
  if (need <= c->hist_cap) return 0;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  drop_graphs(c);
  if (c->hist) HIPCHK(hipFree(c->hist));
  HIPCHK(hipMalloc((void**)&c->hist, need * sizeof(Hist)));
  c->hist_cap = need;
  return 0;
}

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
  auto step = [&](const Args& am, hipStream_t s2, bool first, bool with_a, bool fuse_next) -> int {
    if (with_a) {
      if (first) TRY((col<N, C_A0, true>(am, B, s2)));
      else TRY((col<N, C_A, true>(am, B, s2)));
    }
    TRY((row<N, R_RHS, true>(am, B, s2)));
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
    for (int m = 0; m < maxacc; ++m) {
      Args am = a;
      am.attempt = m;
      TRY(step(am, st, m == 0, m == 0, m + 1 < maxacc));
    }
  }
  TRY(unpack<N>(c->x0, x, B, 0, st));
  TRY(unpack<N>(c->dx0, dx, B, 0, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

// ------------------------------------------------------------------ dispatch
#define DISPATCH_N(n, fn, ...)                            \
  switch (n) {                                            \
    case 16: return fn<16>(__VA_ARGS__);                  \
    case 32: return fn<32>(__VA_ARGS__);                  \
    case 64: return fn<64>(__VA_ARGS__);                  \
    case 128: return fn<128>(__VA_ARGS__);                \
    case 256: return fn<256>(__VA_ARGS__);                \
    default: return fail(1, "unsupported grid size n");   \
  }

template <int N> static int rk4_p(ecs_ctx* c, double* f, int B, const double* dt, int steps, int masked, hipStream_t st) {
  return run_rk4<N, false>(c, f, nullptr, B, dt, nullptr, steps, masked, st);
}
template <int N> static int rk4_t(ecs_ctx* c, double* f, double* df, int B, const double* dt, const double* ddt, int steps, int masked, hipStream_t st) {
  return run_rk4<N, true>(c, f, df, B, dt, ddt, steps, masked, st);
}

extern "C" {

const char* ecs_last_error(void) { return g_err.c_str(); }

int ecs_ctx_create(ecs_ctx** out, int n, double nu, double eta, double b0x, double b0y, const double* fhat_full,
                   int device) {
  if (!out) return fail(1, "null ctx pointer");
  if (n != 16 && n != 32 && n != 64 && n != 128 && n != 256) return fail(1, "unsupported grid size n");
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  HIPCHK(hipSetDevice(device));
  ecs_ctx* c = new ecs_ctx();
  c->n = n; c->dev = device; c->nu = nu; c->eta = eta; c->b0x = b0x; c->b0y = b0y;
  std::vector<double2> tw(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int j = 0; j < n; ++j) {
    const long double ang = -2.0L * pi * (long double)j / (long double)n;
    tw[j] = make_double2((double)cosl(ang), (double)sinl(ang));
  }
  HIPCHK(hipMalloc((void**)&c->tw, n * sizeof(double2)));
  HIPCHK(hipMemcpy(c->tw, tw.data(), n * sizeof(double2), hipMemcpyHostToDevice));
  const int kmax = (n - 1) / 3, KC = kmax + 1, R = 2 * kmax + 1, NH = n / 2 + 1;
  std::vector<double2> fc((size_t)R * KC, make_double2(0.0, 0.0));
  if (fhat_full) {
    for (int r = 0; r < R; ++r) {
      const int x = r <= kmax ? r : r - R + n;
      for (int ky = 0; ky < KC; ++ky) {
        if (r == 0 && ky == 0) continue;  // mask(0,0) = False
        fc[(size_t)r * KC + ky] = make_double2(fhat_full[2 * ((size_t)x * NH + ky)], fhat_full[2 * ((size_t)x * NH + ky) + 1]);
      }
    }
  }
  HIPCHK(hipMalloc((void**)&c->fhat, fc.size() * sizeof(double2)));
  HIPCHK(hipMemcpy(c->fhat, fc.data(), fc.size() * sizeof(double2), hipMemcpyHostToDevice));
  *out = c;
  return 0;
}

int ecs_ctx_destroy(ecs_ctx* c) {
  if (!c) return 0;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  (void)hipSetDevice(c->dev);
  void* ps[] = {c->tw, c->fhat, c->I, c->O, c->x0, c->acc, c->v1, c->v4, c->v5, c->xf, c->out, c->dx0, c->dacc,
                c->dv1, c->dv4, c->dxf, c->dout, c->scal, c->hist, c->s_h, c->s_rs, c->s_rc, c->s_rh, c->s_nacc,
                c->ckpt, c->xtape, c->itape, c->red, c->AO, c->AI, c->xb, c->k1b, c->k2b, c->k3b, c->hb};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  drop_graphs(c);
  if (c->gs) (void)hipStreamDestroy(c->gs);
  if (c->ev_in) (void)hipEventDestroy(c->ev_in);
  if (c->ev_out) (void)hipEventDestroy(c->ev_out);
  if (c->cnt) (void)hipFree(c->cnt);
  if (c->hcnt) (void)hipHostFree(c->hcnt);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->hred) (void)hipHostFree(c->hred);
  delete c;
  (void)hipGetLastError();  // leave no sticky error behind for the caller
  return 0;
}

int ecs_ctx_n(const ecs_ctx* c) { return c ? c->n : 0; }

int ecs_state_vel(ecs_ctx* c, const double* in, double* out, int batch, int diss, void* stream) {
  if (!c || !in || !out || batch < 1) return fail(1, "ecs_state_vel: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  DISPATCH_N(c->n, run_state_vel, c, in, nullptr, out, nullptr, batch, diss, st);
}

int ecs_state_vel_jvp(ecs_ctx* c, const double* in, const double* din, double* out, double* dout, int batch,
                      int diss, void* stream) {
  if (!c || !in || !din || !out || !dout || batch < 1) return fail(1, "ecs_state_vel_jvp: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  DISPATCH_N(c->n, run_state_vel, c, in, din, out, dout, batch, diss, st);
}

int ecs_lawson_rk4(ecs_ctx* c, double* f, int batch, const double* dt, int steps, int masked, void* stream) {
  if (!c || !f || !dt || batch < 1 || steps < 0) return fail(1, "ecs_lawson_rk4: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
DISPATCH_N(c->n, rk4_p, c, f, batch, dt, steps, masked, st);

}

int ecs_lawson_rk4_jvp(ecs_ctx* c, double* f, double* df, int batch, const double* dt, const double* ddt, int steps,
                       int masked, void* stream) {
  if (!c || !f || !df || !dt || !ddt || batch < 1 || steps < 0) return fail(1, "ecs_lawson_rk4_jvp: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
DISPATCH_N(c->n, rk4_t, c, f, df, batch, dt, ddt, steps, masked, st);

}

int ecs_lawson_rk43(ecs_ctx* c, double* x, int batch, const double* T, double h0, double atol, int maxs, int nck,
                    ecs_rk43_info* info, double* hs, int8_t* acc, void* stream) {
  if (!c || !x || !T || batch < 1 || maxs < 1 || nck < 1) return fail(1, "ecs_lawson_rk43: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  DISPATCH_N(c->n, run_rk43, c, x, batch, T, h0, atol, maxs, nck, info, hs, acc, st);
}

int ecs_lawson_rk43_jvp(ecs_ctx* c, double* x, double* dx, int batch, const double* T, const double* dT,
                        const double* hs, const int8_t* acc, const int* attempts, int budget, void* stream) {
  if (!c || !x || !dx || !T || !dT || !hs || !acc || !attempts || batch < 1 || budget < 1)
    return fail(1, "ecs_lawson_rk43_jvp: bad arguments");
  for (int b = 0; b < batch; ++b)
    if (attempts[b] < 0 || attempts[b] > budget) return fail(1, "ecs_lawson_rk43_jvp: attempts out of range");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  DISPATCH_N(c->n, run_rk43_jvp, c, x, dx, batch, T, dT, hs, acc, attempts, budget, st);
}

}  // extern "C"

// ------------------------------------------------------------------ Krylov vectors
static std::mutex g_part_mu;
static std::map<int, double*> g_part;  // per device: 2 * VMAXG partial sums

static int partials(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::recursive_mutex> alk(g_alloc_mu);
  std::lock_guard<std::mutex> lk(g_part_mu);
  auto it = g_part.find(dev);
  if (it == g_part.end()) {
    double* p = nullptr;
    HIPCHK(hipMalloc((void**)&p, 2 * VMAXG * sizeof(double)));
    it = g_part.emplace(dev, p).first;
  }
  *out = it->second;
  return 0;
}

static int red_blocks(int n) { return std::max(1, std::min(VMAXG, (n + 4 * VNT - 1) / (4 * VNT))); }

extern "C" int ecs_dot(const double* a, const double* b, int n, double* out, void* stream) {
  if (!a || !b || !out || n < 1) return fail(1, "ecs_dot: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  double* part = nullptr;
  TRY(partials(&part));
  const int G = red_blocks(n);
  hipLaunchKernelGGL(dot_partial_kernel, dim3(G), dim3(VNT), 0, st, a, b, n, part);
  LAUNCHCHK();
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(64), 0, st, part, G, out);
  LAUNCHCHK();
  return 0;
}

extern "C" int ecs_mgs(double* Q, int ldq, int n, int k, double* w, double* h, void* stream) {
  if (!Q || !w || !h || n < 1 || k < 0 || ldq < n) return fail(1, "ecs_mgs: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  double* part = nullptr;
  TRY(partials(&part));
  const int G = red_blocks(n);
  hipLaunchKernelGGL(dot_partial_kernel, dim3(G), dim3(VNT), 0, st, w, Q, n, part);
  LAUNCHCHK();
  for (int j = 0; j <= k; ++j) {
    double* pin = part + (j & 1) * VMAXG;
    double* pout = part + ((j + 1) & 1) * VMAXG;
    hipLaunchKernelGGL(mgs_update_kernel, dim3(G), dim3(VNT), 0, st, w, Q, ldq, n, j, k, h, pin, pout);
    LAUNCHCHK();
  }
  hipLaunchKernelGGL(mgs_finish_kernel, dim3(G), dim3(VNT), 0, st, w, Q, ldq, n, k, h,
                     part + ((k + 1) & 1) * VMAXG);
  LAUNCHCHK();
  return 0;
}


// ------------------------------------------------------------------ adjoint
template <int N, int MODE>
static int acol(const Args& a, int B, hipStream_t st) {
  constexpr int TY = Tiles<N>::TY;
  dim3 grid((Geo<N>::KC + TY - 1) / TY, 2, B);
  hipLaunchKernelGGL((adj_col_kernel<N, TY, NT, MODE>), grid, dim3(NT), 0, st, a);
  LAUNCHCHK();
  return 0;
}
template <int N>
static int arow(const Args& a, int B, hipStream_t st) {
  constexpr int TX = Tiles<N>::TX;
  dim3 grid(N / TX, 1, B);
  hipLaunchKernelGGL((adj_row_kernel<N, TX, NT>), grid, dim3(NT), 0, st, a);
  LAUNCHCHK();
  return 0;
}

static int grow(double2** p, size_t* cap, size_t need) {
  if (need <= *cap) return 0;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  if (*p) HIPCHK(hipFree(*p));
  *p = nullptr;
  HIPCHK(hipMalloc((void**)p, need * sizeof(double2)));
  *cap = need;
  return 0;
}

static int ensure_adj(ecs_ctx* c, int B) {
  TRY(ensure_cap(c, B));
  if (B <= c->adj_cap) return 0;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  const int nc = c->cap;
  const int KC = (c->n - 1) / 3 + 1;
  const size_t comp = c->compact_elems();
  double2** arrs[] = {&c->xb, &c->k1b, &c->k2b, &c->k3b};
  for (auto p : arrs) {
    if (*p) HIPCHK(hipFree(*p));
    HIPCHK(hipMalloc((void**)p, comp * nc * sizeof(double2)));
  }
  if (c->hb) HIPCHK(hipFree(c->hb));
  HIPCHK(hipMalloc((void**)&c->hb, comp * nc * sizeof(double)));
  if (c->AO) HIPCHK(hipFree(c->AO));
  if (c->AI) HIPCHK(hipFree(c->AI));
  HIPCHK(hipMalloc((void**)&c->AO, (size_t)2 * c->n * KC * nc * sizeof(double2)));
  HIPCHK(hipMalloc((void**)&c->AI, (size_t)8 * c->n * KC * nc * sizeof(double2)));
  c->adj_cap = nc;
  return 0;
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
  const size_t is = (size_t)B * 8 * N * G::KC;
  const size_t need = (size_t)ckpt * 5 * cs;
  if (need > c->vt_cap) {
    if (c->xtape) HIPCHK(hipFree(c->xtape));
    c->xtape = nullptr;
    size_t cap0 = 0;
    TRY(grow(&c->xtape, &cap0, need));
    c->vt_cap = need;
  }
  c->vtape = c->xtape + (size_t)ckpt * cs;
  TRY(grow(&c->itape, &c->it_cap, (size_t)ckpt * 4 * is));
  return 0;
}

// Forward pass of Lawson RK4 storing the compact state at the start of every
// segment of ckpt steps into ck [nseg][B][2][R][KC]; f is advanced in place.
template <int N>
static int run_rk4_ckpt(ecs_ctx* c, double* f, int B, const double* dt, int steps, int masked, int ckpt,
                        double2* ck, hipStream_t st) {
  TRY(ensure_cap(c, B));
  TRY(autotune<N>(c, B, st));
  using G = Geo<N>;
  const size_t cs = (size_t)B * 2 * G::R * G::KC;
  TRY(upload_scalars(c, 0, dt, B, st));
  Args a = base_args(c);
  a.masked_e = masked;
  TRY(pack<N>(f, c->x0, B, st));
  HIPCHK(hipMemcpyAsync(ck, c->x0, cs * sizeof(double2), hipMemcpyDeviceToDevice, st));
  if (steps > 0) TRY((col<N, C_LOAD, false>(a, B, st)));
  for (int s = 0; s < steps; ++s) {
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
                       int masked, int ckpt, hipStream_t st) {
  TRY(adj_buffers<N>(c, B, ckpt));
  TRY(autotune<N>(c, B, st));
  using G = Geo<N>;
  const size_t cs = (size_t)B * 2 * G::R * G::KC;
  const size_t is = (size_t)B * 8 * N * G::KC;
  const int nseg = (steps + ckpt - 1) / ckpt;
  TRY(upload_scalars(c, 0, dt, B, st));
  Args a = base_args(c);
  a.masked_e = masked;
  a.aout = c->AO;
  a.ain = c->AI;
  a.xb = c->xb; a.k1b = c->k1b; a.k2b = c->k2b; a.k3b = c->k3b;
  a.hb = c->hb;
  a.cstride = cs;
  TRY(pack<N>(fbar, c->xb, B, st));
  HIPCHK(hipMemsetAsync(c->hb, 0, cs * sizeof(double), st));
  auto slot = [&](int t, int s) { return c->itape + ((size_t)t * 4 + s) * is; };
  auto xt = [&](int t) { return c->xtape + (size_t)t * cs; };
  auto vt = [&](int t) { return c->vtape + (size_t)t * 4 * cs; };
  for (int j = nseg - 1; j >= 0; --j) {
    const int t0 = j * ckpt;
    const int L = std::min(ckpt, steps - t0);
    HIPCHK(hipMemcpyAsync(c->x0, ck + (size_t)j * cs, cs * sizeof(double2), hipMemcpyDeviceToDevice, st));
    Args f = a;
    f.inter = slot(0, 0);
    TRY((col<N, C_LOAD, false>(f, B, st)));
    for (int t = 0; t < L; ++t) {
      f.vtape = nullptr; f.xtape = nullptr;
      f.inter = slot(t, 0);
      TRY((row<N, R_RHS, false>(f, B, st)));
      f.inter = slot(t, 1); f.vtape = vt(t); f.xtape = xt(t);
      TRY((col<N, C_S1F, false>(f, B, st)));
      f.xtape = nullptr;
      TRY((row<N, R_RHS, false>(f, B, st)));
      f.inter = slot(t, 2); f.vtape = vt(t) + cs;
      TRY((col<N, C_S2, false>(f, B, st)));
      TRY((row<N, R_RHS, false>(f, B, st)));
      f.inter = slot(t, 3); f.vtape = vt(t) + 2 * cs;
      TRY((col<N, C_S3, false>(f, B, st)));
      TRY((row<N, R_RHS, false>(f, B, st)));
      Args f4 = f;
      f4.vtape = vt(t) + 3 * cs;
      f4.inter = (t + 1 < L) ? slot(t + 1, 0) : slot(t, 3);
      f4.skip_inv = (t + 1 == L);
      TRY((col<N, C_S4, false>(f4, B, st)));
    }
    Args b = a;
    b.tx = xt(L - 1);
    b.tv = vt(L - 1);
    TRY((acol<N, B_ENTRY>(b, B, st)));
    for (int t = L - 1; t >= 0; --t) {
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
  }
  TRY(unpack<N>(c->xb, fbar, B, masked ? 1 : 0, st));
  hipLaunchKernelGGL((hb_sum_kernel<N>), dim3(B), dim3(256), 0, st, c->hb, c->red);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(c->hred, c->red, B * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int b2 = 0; b2 < B; ++b2) hbar[b2] = c->hred[b2];
  return 0;
}

static size_t ckpt_elems(int n, int B, int steps, int ckpt) {
  const int kmax = (n - 1) / 3;
  const int nseg = (steps + ckpt - 1) / ckpt;
  return (size_t)std::max(nseg, 1) * B * 2 * (2 * kmax + 1) * (kmax + 1);
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
  int rc = run_rk4_ckpt<N>(c, fwork, B, dt, steps, masked, ckpt, c->ckpt, st);
  if (rc == 0) rc = run_rk4_rev<N>(c, c->ckpt, fbar, hbar, B, dt, steps, masked, ckpt, st);
  HIPCHK(hipFreeAsync(fwork, st));
  return rc;
}

extern "C" int ecs_state_vel_vjp(ecs_ctx* c, const double* in, const double* cot, double* grad, int batch,
                                 int diss, void* stream) {
  if (!c || !in || !cot || !grad || batch < 1) return fail(1, "ecs_state_vel_vjp: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  DISPATCH_N(c->n, run_state_vel_vjp, c, in, cot, grad, batch, diss, st);
}

extern "C" int ecs_lawson_rk4_vjp(ecs_ctx* c, const double* f0, double* fbar, double* hbar, int batch,
                                  const double* dt, int steps, int masked, int ckpt, void* stream) {
  if (!c || !f0 || !fbar || !hbar || !dt || batch < 1 || steps < 1 || ckpt < 1)
    return fail(1, "ecs_lawson_rk4_vjp: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  DISPATCH_N(c->n, run_rk4_vjp, c, f0, fbar, hbar, batch, dt, steps, masked, ckpt, st);
}

extern "C" long long ecs_ckpt_doubles(const ecs_ctx* c, int batch, int steps, int ckpt) {
  if (!c || batch < 1 || steps < 1 || ckpt < 1) return -1;
  return (long long)(2 * ckpt_elems(c->n, batch, steps, ckpt));
}

extern "C" int ecs_lawson_rk4_fwd_ckpt(ecs_ctx* c, double* f, int batch, const double* dt, int steps, int masked,
                                       int ckpt, double* ck, void* stream) {
  if (!c || !f || !dt || !ck || batch < 1 || steps < 1 || ckpt < 1)
    return fail(1, "ecs_lawson_rk4_fwd_ckpt: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  DISPATCH_N(c->n, run_rk4_ckpt, c, f, batch, dt, steps, masked, ckpt, (double2*)ck, st);
}

extern "C" int ecs_lawson_rk4_vjp_ckpt(ecs_ctx* c, const double* ck, double* fbar, double* hbar, int batch,
                                       const double* dt, int steps, int masked, int ckpt, void* stream) {
  if (!c || !ck || !fbar || !hbar || !dt || batch < 1 || steps < 1 || ckpt < 1)
    return fail(1, "ecs_lawson_rk4_vjp_ckpt: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  DISPATCH_N(c->n, run_rk4_rev, c, (const double2*)ck, fbar, hbar, batch, dt, steps, masked, ckpt, st);
}

// ------------------------------------------------------------------ timing hook
// Average duration (ms) of one launch of a fused RHS kernel, timed with HIP
// events on `stream` over `reps` back-to-back launches of the same kernel on
// batch spectra already resident in the context's workspace.
// which: 0 = column kernel (stage 2 glue), 1 = row kernel, 2 = tangent column
// kernel, 3 = tangent row kernel.
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
  auto launch = [&]() -> int {
    switch (which) {
      case 0: return col<N, C_S2, false>(a, B, st);
      case 1: return row<N, R_RHS, false>(a, B, st);
      case 2: return col<N, C_S2, true>(a, B, st);
      default: return row<N, R_RHS, true>(a, B, st);
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

extern "C" int ecs_time_kernel(ecs_ctx* c, int batch, int which, int reps, double* ms, void* stream) {
  if (!c || !ms || batch < 1 || reps < 1 || which < 0 || which > 3) return fail(1, "ecs_time_kernel: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  DISPATCH_N(c->n, run_time_kernel, c, batch, which, reps, ms, st);
}

extern "C" int ecs_tuned_tiles(int n, int batch, int* out) {
  if (!out || batch < 1) return fail(1, "ecs_tuned_tiles: bad arguments");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int kind = 0; kind < 4; ++kind) {
    auto it = g_tune.find(tune_key(n, kind, batch));
    const int v = it == g_tune.end() ? -1 : it->second;
    out[kind] = v < 0 ? -1 : (1 << v);
  }
  return 0;
}
