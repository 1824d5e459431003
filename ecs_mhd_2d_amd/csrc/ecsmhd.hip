// Host orchestration and C-ABI of the fused MHD kernels (gfx950, FP64):
// error reporting, context life cycle, workspace, autotuner tables, Krylov
// vector kernels, and the entry points, which dispatch to the per-grid-size
// drivers compiled in grid_n*.hip.
#include "host.h"
#include "vec_kernels.h"
#include "recurrence_kernels.h"

static thread_local std::string g_err;

// Workspace of ecs_state_vel_noalloc: one copy per GPU, held until the process
// ends.  Host code never uses this name as an address: each context keeps the
// device address hipGetSymbolAddress gives on its GPU (ctx->noalloc).
static __device__ __align__(256) double2 scratch[ECS_NOALLOC_SCRATCH_BYTES / sizeof(double2)];
static std::mutex g_scratch_mu;
static std::map<int, double2*> g_scratch_addr;  // device id -> address of `scratch` there

// Address of `scratch` on the current device (looked up once per device).
static int scratch_address(int dev, double2** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_addr.find(dev);
  if (it == g_scratch_addr.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(2, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    it = g_scratch_addr.emplace(dev, (double2*)p).first;
  }
  *out = it->second;
  return 0;
}
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

std::recursive_mutex g_alloc_mu;

void drop_graphs(ecs_ctx* c) {
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
}

int ensure_cap(ecs_ctx* c, int B) {
  if (B <= c->cap) return 0;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  drop_graphs(c);
  int nc = std::max(B, c->cap * 2);
  const size_t comp = c->compact_elems();
  const size_t inter = (size_t)16 * c->n * kp_of(c->n);
  const size_t rowo = (size_t)4 * c->n * kp_of(c->n);
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
  if (c->nstep) HIPCHK(hipFree(c->nstep));
  if (c->hnstep) HIPCHK(hipHostFree(c->hnstep));
  HIPCHK(hipMalloc((void**)&c->nstep, nc * sizeof(int)));
  HIPCHK(hipHostMalloc((void**)&c->hnstep, nc * sizeof(int), hipHostMallocDefault));
  if (c->red) HIPCHK(hipFree(c->red));
  if (c->hred) HIPCHK(hipHostFree(c->hred));
  HIPCHK(hipMalloc((void**)&c->red, (size_t)64 * nc * sizeof(double)));
  HIPCHK(hipHostMalloc((void**)&c->hred, (size_t)64 * nc * sizeof(double), hipHostMallocDefault));
  c->cap = nc;
  return 0;
}

Args base_args(ecs_ctx* c) {
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

// ------------------------------------------------------------------ autotuner tables
std::mutex g_tune_mu;
std::map<long long, int> g_tune;
std::map<long long, int> g_plan;

static int bucket_of(int B) {
  int k = 0;
  while ((1 << k) < B && k < 12) ++k;
  return k;
}
long long tune_key(int n, int kind, int B) { return ((long long)n * 8 + kind) * 64 + bucket_of(B); }
long long plan_key(int n, int kind) { return ((long long)n * 8 + kind) * 64 + 63; }

int plan_table(int n, int kind) {
  // families: column, tangent column, row, tangent row
  static const int kPlanTable[5][4] = {
      {1, 1, 1, 0},   // N = 16
      {0, 0, 0, 0},   // N = 32
      {1, 1, 0, 0},   // N = 64
      {1, 1, 1, 0},   // N = 128
      {1, 1, 1, 0},   // N = 256
  };
  const int row = n == 16 ? 0 : n == 32 ? 1 : n == 64 ? 2 : n == 128 ? 3 : 4;
  return kPlanTable[row][kind & 3];
}

int tuned(int n, int kind, int B, int dflt) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(tune_key(n, kind, B));
  if (it != g_tune.end()) return it->second;
  // an untuned bucket keeps the default tile within the grid size's fixed plan
  auto ip = g_plan.find(plan_key(n, base_kind(kind)));
  const int plan = ip == g_plan.end() ? plan_table(n, base_kind(kind)) : ip->second;
  return 3 * plan + dflt % 3;
}

int plan_of(int n, int kind) {
  kind = base_kind(kind);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto ip = g_plan.find(plan_key(n, kind));
  return ip == g_plan.end() ? plan_table(n, kind) : ip->second;
}

int upload_scalars(ecs_ctx* c, int slot, const double* vals, int B, hipStream_t st) {
  // the previous copy out of this pinned slot may still be queued behind
  // earlier work: wait for it before overwriting the staging memory
  if (!c->ev_scal[slot]) HIPCHK(hipEventCreateWithFlags(&c->ev_scal[slot], hipEventDisableTiming));
  else HIPCHK(hipEventSynchronize(c->ev_scal[slot]));
  double* h = c->hscal + (size_t)slot * c->cap;
  for (int b = 0; b < B; ++b) h[b] = vals[b];
  HIPCHK(hipMemcpyAsync(c->scal + (size_t)slot * c->cap, h, B * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(c->ev_scal[slot], st));
  return 0;
}

int upload_ints(ecs_ctx* c, const int* vals, int B, hipStream_t st) {
  if (!c->ev_nstep) HIPCHK(hipEventCreateWithFlags(&c->ev_nstep, hipEventDisableTiming));
  else HIPCHK(hipEventSynchronize(c->ev_nstep));
  for (int b = 0; b < B; ++b) c->hnstep[b] = vals[b];
  HIPCHK(hipMemcpyAsync(c->nstep, c->hnstep, B * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(c->ev_nstep, st));
  return 0;
}


// ------------------------------------------------------------------ hipGraph chunks
bool use_graphs() { return getenv("ECSMHD_GRAPH") != nullptr; }

std::string graph_key(const char* tag, const Args& a, int B, int K) {
  std::string k(tag);
  k.append((const char*)&a, sizeof(Args));
  k.append((const char*)&B, sizeof(int));
  k.append((const char*)&K, sizeof(int));
  return k;
}

int set_counter(ecs_ctx* c, int v, hipStream_t st) {
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

int ensure_hist(ecs_ctx* c, size_t need) {
  if (need <= c->hist_cap) return 0;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  drop_graphs(c);
  if (c->hist) HIPCHK(hipFree(c->hist));
  HIPCHK(hipMalloc((void**)&c->hist, need * sizeof(Hist)));
  c->hist_cap = need;
  return 0;
}


int grow(double2** p, size_t* cap, size_t need) {
  if (need <= *cap) return 0;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  if (*p) HIPCHK(hipFree(*p));
  *p = nullptr;
  HIPCHK(hipMalloc((void**)p, need * sizeof(double2)));
  *cap = need;
  return 0;
}

int ensure_adj(ecs_ctx* c, int B) {
  TRY(ensure_cap(c, B));
  if (B <= c->adj_cap) return 0;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  const int nc = c->cap;
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
  HIPCHK(hipMalloc((void**)&c->AO, (size_t)2 * c->n * kp_of(c->n) * nc * sizeof(double2)));
  HIPCHK(hipMalloc((void**)&c->AI, (size_t)8 * c->n * kp_of(c->n) * nc * sizeof(double2)));
  c->adj_cap = nc;
  return 0;
}

size_t ckpt_elems(int n, int B, int steps, int ckpt) {
  const int kmax = (n - 1) / 3;
  const int nseg = (steps + ckpt - 1) / ckpt;
  return (size_t)std::max(nseg, 1) * B * 2 * (2 * kmax + 1) * (kmax + 1);
}


const GridOps* grid_ops(int n) {
  switch (n) {
    case 16: return grid_ops_16();
    case 32: return grid_ops_32();
    case 64: return grid_ops_64();
    case 128: return grid_ops_128();
    case 256: return grid_ops_256();
    default: return nullptr;
  }
}
// contexts exist only for supported sizes (ecs_ctx_create checks), so the table is never null here
#define OPS(n) grid_ops(n)

// Twiddles and masked forcing spectrum of a new context (device memory).
static int ctx_tables(ecs_ctx* c, const double* fhat_full) {
  const int n = c->n;
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
  return 0;
}

extern "C" {

const char* ecs_last_error(void) { return g_err.c_str(); }

int ecs_ctx_create(ecs_ctx** out, int n, double nu, double eta, double b0x, double b0y, const double* fhat_full,
                   int device) {
  if (!out) return fail(1, "null ctx pointer");
  if (n != 16 && n != 32 && n != 64 && n != 128 && n != 256) return fail(1, "unsupported grid size n");
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  DevGuard dg(device);
  if (!dg.ok) return fail(2, "ecs_ctx_create: cannot make the device current");
  ecs_ctx* c = new ecs_ctx();
  c->n = n; c->dev = device; c->nu = nu; c->eta = eta; c->b0x = b0x; c->b0y = b0y;
  int rc = ctx_tables(c, fhat_full);
  if (!rc) rc = scratch_address(device, &c->noalloc);
  if (rc) {
    const std::string msg = g_err;
    ecs_ctx_destroy(c);  // frees what was allocated before the failure
    g_err = msg;
    return rc;
  }
  *out = c;
  return 0;
}


int ecs_ctx_destroy(ecs_ctx* c) {
  if (!c) return 0;
  std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
  DevGuard dg(c->dev);
  void* ps[] = {c->tw, c->fhat, c->I, c->O, c->x0, c->acc, c->v1, c->v4, c->v5, c->xf, c->out, c->dx0, c->dacc,
                c->dv1, c->dv4, c->dxf, c->dout, c->scal, c->hist, c->s_h, c->s_rs, c->s_rc, c->s_rh, c->s_nacc,
                c->ckpt, c->xtape, c->itape, c->xtape2, c->itape2, c->red, c->AO, c->AI, c->xb, c->k1b, c->k2b,
                c->k3b, c->hb, c->gen, c->tab, c->gcomb};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  drop_graphs(c);
  if (c->gs) (void)hipStreamDestroy(c->gs);
  if (c->side) (void)hipStreamDestroy(c->side);
  for (hipEvent_t e : c->ev_rev)
    if (e) (void)hipEventDestroy(e);
  if (c->ev_in) (void)hipEventDestroy(c->ev_in);
  if (c->ev_out) (void)hipEventDestroy(c->ev_out);
  if (c->cnt) (void)hipFree(c->cnt);
  if (c->hcnt) (void)hipHostFree(c->hcnt);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->hred) (void)hipHostFree(c->hred);
  if (c->nstep) (void)hipFree(c->nstep);
  if (c->hnstep) (void)hipHostFree(c->hnstep);
  if (c->ev_nstep) (void)hipEventDestroy(c->ev_nstep);
  for (hipEvent_t e : c->probe_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->ev_scal)
    if (e) (void)hipEventDestroy(e);
  delete c;
  (void)hipGetLastError();  // leave no sticky error behind for the caller
  return 0;
}

int ecs_ctx_n(const ecs_ctx* c) { return c ? c->n : 0; }

extern "C" int ecs_ctx_set_rev_overlap(ecs_ctx* c, int enable) {
  if (!c) return fail(1, "ecs_ctx_set_rev_overlap: null context");
  c->rev_overlap = enable != 0;
  return 0;
}

int ecs_state_vel(ecs_ctx* c, const double* in, double* out, int batch, int diss, void* stream) {
  if (!c || !in || !out || batch < 1) return fail(1, "ecs_state_vel: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->state_vel(c, in, nullptr, out, nullptr, batch, diss, st);
}

int ecs_state_vel_noalloc(ecs_ctx* c, const double* in, double* out, int batch, int diss, void* stream) {
  if (!c || !in || !out || batch < 1) return fail(1, "ecs_state_vel_noalloc: bad arguments");
  GUARD(c);
  if (!c->noalloc) return fail(2, "ecs_state_vel_noalloc: no scratch address for the context's device");
  return OPS(c->n)->state_vel_noalloc(c, in, out, batch, diss, c->noalloc,
                                      ECS_NOALLOC_SCRATCH_BYTES / sizeof(double2), (hipStream_t)stream);
}

int ecs_state_vel_jvp(ecs_ctx* c, const double* in, const double* din, double* out, double* dout, int batch,
                      int diss, void* stream) {
  if (!c || !in || !din || !out || !dout || batch < 1) return fail(1, "ecs_state_vel_jvp: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->state_vel(c, in, din, out, dout, batch, diss, st);
}

int ecs_lawson_rk4(ecs_ctx* c, double* f, int batch, const double* dt, int steps, int masked, void* stream) {
  if (!c || !f || !dt || batch < 1 || steps < 0) return fail(1, "ecs_lawson_rk4: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->rk4(c, f, batch, dt, steps, masked, st);

}

int ecs_lawson_rk4_jvp(ecs_ctx* c, double* f, double* df, int batch, const double* dt, const double* ddt, int steps,
                       int masked, void* stream) {
  if (!c || !f || !df || !dt || !ddt || batch < 1 || steps < 0) return fail(1, "ecs_lawson_rk4_jvp: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->rk4_jvp(c, f, df, batch, dt, ddt, steps, masked, st);

}

int ecs_lawson_rk43(ecs_ctx* c, double* x, int batch, const double* T, double h0, double atol, int maxs, int nck,
                    ecs_rk43_info* info, double* hs, int8_t* acc, void* stream) {
  if (!c || !x || !T || batch < 1 || maxs < 1 || nck < 1) return fail(1, "ecs_lawson_rk43: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->rk43(c, x, batch, T, h0, atol, maxs, nck, info, hs, acc, st);
}

int ecs_lawson_rk43_jvp(ecs_ctx* c, double* x, double* dx, int batch, const double* T, const double* dT,
                        const double* hs, const int8_t* acc, const int* attempts, int budget, void* stream) {
  if (!c || !x || !dx || !T || !dT || !hs || !acc || !attempts || batch < 1 || budget < 1)
    return fail(1, "ecs_lawson_rk43_jvp: bad arguments");
  for (int b = 0; b < batch; ++b)
    if (attempts[b] < 0 || attempts[b] > budget) return fail(1, "ecs_lawson_rk43_jvp: attempts out of range");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->rk43_jvp(c, x, dx, batch, T, dT, hs, acc, attempts, budget, st);
}

int ecs_rk4(ecs_ctx* c, double* f, int batch, const double* dt, int steps, void* stream) {
  if (!c || !f || !dt || batch < 1 || steps < 0) return fail(1, "ecs_rk4: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen(c, f, batch, dt, steps, 0, 0, (hipStream_t)stream);
}

int ecs_rk4_jvp(ecs_ctx* c, double* f, double* df, int batch, const double* dt, const double* ddt, int steps,
                void* stream) {
  if (!c || !f || !df || !dt || !ddt || batch < 1 || steps < 0) return fail(1, "ecs_rk4_jvp: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen_jvp(c, f, df, batch, dt, ddt, steps, 0, 0, (hipStream_t)stream);
}

int ecs_lawson_rk6(ecs_ctx* c, double* f, int batch, const double* dt, int steps, int masked, void* stream) {
  if (!c || !f || !dt || batch < 1 || steps < 0) return fail(1, "ecs_lawson_rk6: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen(c, f, batch, dt, steps, 1, masked ? 1 : 0, (hipStream_t)stream);
}

int ecs_lawson_rk6_jvp(ecs_ctx* c, double* f, double* df, int batch, const double* dt, const double* ddt, int steps,
                       int masked, void* stream) {
  if (!c || !f || !df || !dt || !ddt || batch < 1 || steps < 0) return fail(1, "ecs_lawson_rk6_jvp: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen_jvp(c, f, df, batch, dt, ddt, steps, 1, masked ? 1 : 0, (hipStream_t)stream);
}

int ecs_tdrk4(ecs_ctx* c, double* f, int batch, const double* dt, int steps, void* stream) {
  if (!c || !f || !dt || batch < 1 || steps < 0) return fail(1, "ecs_tdrk4: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen(c, f, batch, dt, steps, 2, 0, (hipStream_t)stream);
}

int ecs_tdrk4_jvp(ecs_ctx* c, double* f, double* df, int batch, const double* dt, const double* ddt, int steps,
                  void* stream) {
  if (!c || !f || !df || !dt || !ddt || batch < 1 || steps < 0) return fail(1, "ecs_tdrk4_jvp: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen_jvp(c, f, df, batch, dt, ddt, steps, 2, 0, (hipStream_t)stream);
}

int ecs_rk4_vjp(ecs_ctx* c, const double* f0, double* fbar, double* hbar, int batch, const double* dt, int steps,
                int ckpt, void* stream) {
  if (!c || !f0 || !fbar || !hbar || !dt || batch < 1 || steps < 1 || ckpt < 1)
    return fail(1, "ecs_rk4_vjp: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen_vjp(c, f0, fbar, hbar, batch, dt, steps, 0, 0, ckpt, nullptr, nullptr, nullptr,
                            (hipStream_t)stream);
}

int ecs_lawson_rk6_vjp(ecs_ctx* c, const double* f0, double* fbar, double* hbar, int batch, const double* dt,
                       int steps, int masked, int ckpt, void* stream) {
  if (!c || !f0 || !fbar || !hbar || !dt || batch < 1 || steps < 1 || ckpt < 1)
    return fail(1, "ecs_lawson_rk6_vjp: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen_vjp(c, f0, fbar, hbar, batch, dt, steps, 1, masked ? 1 : 0, ckpt, nullptr, nullptr, nullptr,
                            (hipStream_t)stream);
}

int ecs_tdrk4_vjp(ecs_ctx* c, const double* f0, double* fbar, double* hbar, int batch, const double* dt, int steps,
                  int ckpt, void* stream) {
  if (!c || !f0 || !fbar || !hbar || !dt || batch < 1 || steps < 1 || ckpt < 1)
    return fail(1, "ecs_tdrk4_vjp: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen_vjp(c, f0, fbar, hbar, batch, dt, steps, 2, 0, ckpt, nullptr, nullptr, nullptr,
                            (hipStream_t)stream);
}

int ecs_lawson_rk43_vjp(ecs_ctx* c, const double* x0, double* xbar, double* Tbar, int batch, const double* T,
                        const double* hs, const int8_t* acc, const int* attempts, int budget, int ckpt,
                        void* stream) {
  if (!c || !x0 || !xbar || !Tbar || !T || !hs || !acc || !attempts || batch < 1 || budget < 1 || ckpt < 1)
    return fail(1, "ecs_lawson_rk43_vjp: bad arguments");
  for (int b = 0; b < batch; ++b)
    if (attempts[b] < 0 || attempts[b] > budget) return fail(1, "ecs_lawson_rk43_vjp: attempts out of range");
  GUARD(c);
  return OPS(c->n)->rk43_vjp(c, x0, xbar, Tbar, batch, T, hs, acc, attempts, budget, ckpt, nullptr, nullptr,
                             (hipStream_t)stream);
}

int ecs_gen_fwd_ckpt(ecs_ctx* c, double* f, int batch, const double* dt, int steps, int scheme, int masked, int ckpt,
                     double* ck, void* stream) {
  if (!c || !f || !dt || !ck || batch < 1 || steps < 1 || ckpt < 1 || scheme < 0 || scheme > 2)
    return fail(1, "ecs_gen_fwd_ckpt: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen_vjp(c, f, nullptr, nullptr, batch, dt, steps, scheme, scheme == 1 && masked ? 1 : 0, ckpt,
                            nullptr, (double2*)ck, f, (hipStream_t)stream);
}

int ecs_gen_vjp_ckpt(ecs_ctx* c, const double* ck, double* fbar, double* hbar, int batch, const double* dt, int steps,
                     int scheme, int masked, int ckpt, void* stream) {
  if (!c || !ck || !fbar || !hbar || !dt || batch < 1 || steps < 1 || ckpt < 1 || scheme < 0 || scheme > 2)
    return fail(1, "ecs_gen_vjp_ckpt: bad arguments");
  GUARD(c);
  return OPS(c->n)->gen_vjp(c, nullptr, fbar, hbar, batch, dt, steps, scheme, scheme == 1 && masked ? 1 : 0, ckpt,
                            (const double2*)ck, nullptr, nullptr, (hipStream_t)stream);
}

static int check_rk43(const char* who, const int* attempts, int batch, int budget) {
  for (int b = 0; b < batch; ++b)
    if (attempts[b] < 0 || attempts[b] > budget) return fail(1, std::string(who) + ": attempts out of range");
  return 0;
}

int ecs_lawson_rk43_fwd_ckpt(ecs_ctx* c, const double* x0, int batch, const double* T, const double* hs,
                             const int8_t* acc, const int* attempts, int budget, int ckpt, double* ck, void* stream) {
  if (!c || !x0 || !T || !hs || !acc || !attempts || !ck || batch < 1 || budget < 1 || ckpt < 1)
    return fail(1, "ecs_lawson_rk43_fwd_ckpt: bad arguments");
  TRY(check_rk43("ecs_lawson_rk43_fwd_ckpt", attempts, batch, budget));
  GUARD(c);
  return OPS(c->n)->rk43_vjp(c, x0, nullptr, nullptr, batch, T, hs, acc, attempts, budget, ckpt, nullptr,
                             (double2*)ck, (hipStream_t)stream);
}

int ecs_lawson_rk43_vjp_ckpt(ecs_ctx* c, const double* ck, double* xbar, double* Tbar, int batch, const double* T,
                             const double* hs, const int8_t* acc, const int* attempts, int budget, int ckpt,
                             void* stream) {
  if (!c || !ck || !xbar || !Tbar || !T || !hs || !acc || !attempts || batch < 1 || budget < 1 || ckpt < 1)
    return fail(1, "ecs_lawson_rk43_vjp_ckpt: bad arguments");
  TRY(check_rk43("ecs_lawson_rk43_vjp_ckpt", attempts, batch, budget));
  GUARD(c);
  return OPS(c->n)->rk43_vjp(c, nullptr, xbar, Tbar, batch, T, hs, acc, attempts, budget, ckpt, (const double2*)ck,
                             nullptr, (hipStream_t)stream);
}

}  // extern "C"

// ------------------------------------------------------------------ Krylov vectors
// Partial sums of the two-level reductions, one buffer per (device, stream)
// so that calls on different streams never share one (2 * VMAXG doubles,
// kept until the process ends).
static std::mutex g_part_mu;
static std::map<std::pair<int, hipStream_t>, double*> g_part;

static int partials(double** out, hipStream_t st) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::recursive_mutex> alk(g_alloc_mu);
  std::lock_guard<std::mutex> lk(g_part_mu);
  auto it = g_part.find({dev, st});
  if (it == g_part.end()) {
    double* p = nullptr;
    HIPCHK(hipMalloc((void**)&p, 2 * VMAXG * sizeof(double)));
    it = g_part.emplace(std::make_pair(dev, st), p).first;
  }
  *out = it->second;
  return 0;
}

static int red_blocks(int n) { return std::max(1, std::min(VMAXG, (n + 4 * VNT - 1) / (4 * VNT))); }

extern "C" int ecs_dot(const double* a, const double* b, int n, double* out, void* stream) {
  if (!a || !b || !out || n < 1) return fail(1, "ecs_dot: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  double* part = nullptr;
  TRY(partials(&part, st));
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
  TRY(partials(&part, st));
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

// Scratch for the discarded dots of reorthogonalisation passes, per (device, stream).
static std::map<std::pair<int, hipStream_t>, std::pair<double*, int>> g_hscr;

extern "C" int ecs_mgs_reorth(double* Q, int ldq, int n, int k, double* w, double* h, int passes, void* stream) {
  if (!Q || !w || !h || n < 1 || k < 0 || ldq < n || passes < 0) return fail(1, "ecs_mgs_reorth: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  double* part = nullptr;
  TRY(partials(&part, st));
  double* scr = nullptr;
  {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::recursive_mutex> alk(g_alloc_mu);
    std::lock_guard<std::mutex> lk(g_part_mu);
    auto& e = g_hscr[std::make_pair(dev, st)];
    if (e.second < k + 2) {
      if (e.first) HIPCHK(hipFree(e.first));
      e.first = nullptr;
      e.second = 0;
      HIPCHK(hipMalloc((void**)&e.first, (size_t)(k + 2) * sizeof(double)));
      e.second = k + 2;
    }
    scr = e.first;
  }
  const int G = red_blocks(n);
  int slot = 0;
  for (int p = 0; p <= passes; ++p) {
    double* hp = p == 0 ? h : scr;   // only the first pass's dots enter H
    hipLaunchKernelGGL(dot_partial_kernel, dim3(G), dim3(VNT), 0, st, w, Q, n, part + slot * VMAXG);
    LAUNCHCHK();
    for (int j = 0; j <= k; ++j) {
      double* pin = part + slot * VMAXG;
      slot ^= 1;
      double* pout = part + slot * VMAXG;
      hipLaunchKernelGGL(mgs_update_kernel, dim3(G), dim3(VNT), 0, st, w, Q, ldq, n, j, k, hp, pin, pout);
      LAUNCHCHK();
    }
  }
  hipLaunchKernelGGL(mgs_finish_kernel, dim3(G), dim3(VNT), 0, st, w, Q, ldq, n, k, h, part + slot * VMAXG);
  LAUNCHCHK();
  return 0;
}



extern "C" int ecs_recurrence(const double* fs, int nsnap, long long modes, double* dist, void* stream) {
  if (!fs || !dist || nsnap < 1 || modes < 1) return fail(1, "ecs_recurrence: bad arguments");
  const int nt = (nsnap + REC_TS - 1) / REC_TS;
  hipLaunchKernelGGL(recurrence_kernel, dim3(nt, nt), dim3(256), 0, (hipStream_t)stream, (const double2*)fs, nsnap,
                     modes, dist);
  LAUNCHCHK();
  return 0;
}

extern "C" int ecs_state_vel_vjp(ecs_ctx* c, const double* in, const double* cot, double* grad, int batch,
                                 int diss, void* stream) {
  if (!c || !in || !cot || !grad || batch < 1) return fail(1, "ecs_state_vel_vjp: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->state_vel_vjp(c, in, cot, grad, batch, diss, st);
}

extern "C" int ecs_lawson_rk4_vjp(ecs_ctx* c, const double* f0, double* fbar, double* hbar, int batch,
                                  const double* dt, int steps, int masked, int ckpt, void* stream) {
  if (!c || !f0 || !fbar || !hbar || !dt || batch < 1 || steps < 1 || ckpt < 1)
    return fail(1, "ecs_lawson_rk4_vjp: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->rk4_vjp(c, f0, fbar, hbar, batch, dt, steps, masked, ckpt, st);
}

extern "C" long long ecs_ckpt_doubles(const ecs_ctx* c, int batch, int steps, int ckpt) {
  if (!c || batch < 1 || steps < 1 || ckpt < 1) return -1;
  return (long long)(2 * ckpt_elems(c->n, batch, steps, ckpt));
}

// nsteps (host [batch], optional): per-candidate step counts, each in
// [1, steps]; in the reverse pass each must be a multiple of ckpt or equal
// steps, so that a candidate is active for whole segments only.
static int check_nsteps(const char* who, const int* nsteps, int batch, int steps, int ckpt) {
  if (!nsteps) return 0;
  for (int b = 0; b < batch; ++b) {
    const int k = nsteps[b];
    if (k < 1 || k > steps || (k != steps && k % ckpt != 0))
      return fail(1, std::string(who) + ": nsteps[b] must be in [1, steps] and a multiple of ckpt or equal to steps");
  }
  return 0;
}

extern "C" int ecs_lawson_rk4_fwd_ckpt(ecs_ctx* c, double* f, int batch, const double* dt, int steps, int masked,
                                       int ckpt, double* ck, const int* nsteps, void* stream) {
  if (!c || !f || !dt || !ck || batch < 1 || steps < 1 || ckpt < 1)
    return fail(1, "ecs_lawson_rk4_fwd_ckpt: bad arguments");
  TRY(check_nsteps("ecs_lawson_rk4_fwd_ckpt", nsteps, batch, steps, ckpt));
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->rk4_ckpt(c, f, batch, dt, steps, masked, ckpt, (double2*)ck, nsteps, st);
}

extern "C" int ecs_lawson_rk4_vjp_ckpt(ecs_ctx* c, const double* ck, double* fbar, double* hbar, int batch,
                                       const double* dt, int steps, int masked, int ckpt, const int* nsteps,
                                       void* stream) {
  if (!c || !ck || !fbar || !hbar || !dt || batch < 1 || steps < 1 || ckpt < 1)
    return fail(1, "ecs_lawson_rk4_vjp_ckpt: bad arguments");
  TRY(check_nsteps("ecs_lawson_rk4_vjp_ckpt", nsteps, batch, steps, ckpt));
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->rk4_rev(c, (const double2*)ck, fbar, hbar, batch, dt, steps, masked, ckpt, nsteps, st);
}

extern "C" int ecs_time_kernel(ecs_ctx* c, int batch, int which, int reps, double* ms, void* stream) {
  if (!c || !ms || batch < 1 || reps < 1 || which < 0 || which > 5) return fail(1, "ecs_time_kernel: bad arguments");
  GUARD(c);
  hipStream_t st = (hipStream_t)stream;
  return OPS(c->n)->time_kernel(c, batch, which, reps, ms, st);
}

extern "C" int ecs_replay_probe(ecs_ctx* c, int enable, int every, double* avg_ms, int* samples) {
  if (!c || (enable && every < 1)) return fail(1, "ecs_replay_probe: bad arguments");
  GUARD(c);
  constexpr size_t kPairs = 1024;
  if (enable) {
    if (c->probe_ev.empty()) {
      c->probe_ev.resize(2 * kPairs, nullptr);
      for (auto& e : c->probe_ev) HIPCHK(hipEventCreate(&e));
    }
    c->probe_every = every;
    c->probe_used = 0;
    return 0;
  }
  double sum = 0.0;
  const size_t used = c->probe_used;
  if (used) HIPCHK(hipEventSynchronize(c->probe_ev[used - 1]));
  for (size_t i = 0; i + 1 < used; i += 2) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->probe_ev[i], c->probe_ev[i + 1]));
    sum += t;
  }
  if (avg_ms) *avg_ms = used ? sum / (double)(used / 2) : 0.0;
  if (samples) *samples = (int)(used / 2);
  c->probe_every = 0;
  c->probe_used = 0;
  return 0;
}

extern "C" int ecs_tuned_tiles(int n, int batch, int* out) {
  if (!out || batch < 1) return fail(1, "ecs_tuned_tiles: bad arguments");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int kind = 0; kind < 4; ++kind) {
    auto it = g_tune.find(tune_key(n, kind, batch));
    const int v = it == g_tune.end() ? -1 : it->second;
    out[kind] = v < 0 ? -1 : var_tile(v);
  }
  return 0;
}

extern "C" int ecs_tuned_variants(int n, int batch, int* out) {
  if (!out || batch < 1) return fail(1, "ecs_tuned_variants: bad arguments");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int kind = 0; kind < 4; ++kind) {
    auto it = g_tune.find(tune_key(n, kind, batch));
    out[kind] = it == g_tune.end() ? -1 : it->second;
  }
  return 0;
}

extern "C" int ecs_tuned_replay_variants(int n, int batch, int* out) {
  if (!out || batch < 1) return fail(1, "ecs_tuned_replay_variants: bad arguments");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int k = 0; k < 2; ++k) {
    auto it = g_tune.find(tune_key(n, k == 0 ? K_COLR : K_ROWR, batch));
    out[k] = it == g_tune.end() ? -1 : it->second;
  }
  return 0;
}

extern "C" int ecs_tuned_splits(int n, int batch, int* out) {
  if (!out || batch < 1) return fail(1, "ecs_tuned_splits: bad arguments");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int kind = 0; kind < 2; ++kind) {
    auto it = g_tune.find(tune_key(n, kind, batch));
    out[kind] = it == g_tune.end() ? -1 : var_split(it->second);
  }
  return 0;
}

extern "C" int ecs_tuned_plans(int n, int* out) {
  if (!out) return fail(1, "ecs_tuned_plans: bad arguments");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int kind = 0; kind < 4; ++kind) {
    auto it = g_plan.find(plan_key(n, kind));
    out[kind] = it == g_plan.end() ? -1 : (it->second == 0 ? 16 : 8);
  }
  return 0;
}
