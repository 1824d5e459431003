/* C-ABI of the MI355X-native periodic-orbit hot path of 2D MHD.
 *
 * All arrays are caller-owned device buffers, C-order, FP64.  A "full
 * spectrum" is complex128 [batch][2][N][N/2+1] (interleaved re, im doubles),
 * the layout of jnp.fft.rfft2 of real fields [batch][2][N][N] with field 0 =
 * vorticity omega and field 1 = current j (mhd_jax.py:105-120).  Every call
 * is asynchronous on `stream` (a hipStream_t, 0 = default) unless stated.
 * Return value: 0 on success, otherwise a nonzero code; ecs_last_error()
 * describes the last failure of the calling thread.
 *
 * The reference (mgolden30/ECS_MHD_2D) has no native boundary: its operator
 * API is a set of JAX functions.  Each entry point below replaces the
 * reference function it names.
 */
#ifndef ECSMHD_H
#define ECSMHD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ecs_ctx ecs_ctx;

typedef struct ecs_rk43_info {
  int completed;  /* |s - 1| < 1e-10 */
  int accepted;
  int rejected;
  int fevals;     /* 1 + 4 (accepted + rejected) */
  int attempts;   /* accepted + rejected */
  int pad;
  double s;       /* pseudo-time reached */
} ecs_rk43_info;

/* Domain tables and physics of one grid size.  Replaces construct_domain
 * (jax_scripts/lib/mhd_jax.py:20-94) plus the nu, eta, b0, forcing entries of
 * param_dict.  fhat_full: host pointer to the complex128 [N][N/2+1] rfft2 of
 * the real-space forcing (it is masked inside).  n in {16,32,64,128,256}. */
int ecs_ctx_create(ecs_ctx** out, int n, double nu, double eta, double b0x, double b0y,
                   const double* fhat_full, int device);
int ecs_ctx_destroy(ecs_ctx* ctx);
const char* ecs_last_error(void);
int ecs_ctx_n(const ecs_ctx* ctx);

/* Execution option of a context (no counterpart in the reference, whose
 * reverse mode is XLA's): with enable = 1 (the default) the reverse passes
 * (Lawson RK4, classical RK4, Lawson RK6, TDRK4, RK4(3)) recompute the next
 * segment on a stream of the context while the caller's stream sweeps the
 * current one backward (two step tapes, each when it fits 4 GiB); 0 keeps one
 * stream.  The results are the same bits either way; callers that already run
 * several contexts concurrently (a serving loop) may prefer one stream per
 * call. */
int ecs_ctx_set_rev_overlap(ecs_ctx* ctx, int enable);

/* state_vel(fields, param_dict, include_dissipation)  (mhd_jax.py:105-175).
 * in, out: full spectra [batch][2][N][N/2+1]. */
int ecs_state_vel(ecs_ctx* ctx, const double* in, double* out, int batch, int include_dissipation,
                  void* stream);

/* Allocation-free variant of ecs_state_vel: the same kernels and the same
 * result.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its workspace is a static device array of
 * ECS_NOALLOC_SCRATCH_BYTES that the library holds on each GPU until the
 * process ends; the call allocates nothing and needs no earlier call.
 * A batch needs batch * (2 C + 10 N KP) * 16 bytes, C = 2 (2K+1)(K+1),
 * K = (N-1)/3, KP = K+1 rounded up to a multiple of 8: at most 14 candidates
 * at N=256, 55 at N=128; a larger batch returns 1.  Every call on one GPU
 * shares the array: calls that overlap in time on one GPU (on other streams
 * or other contexts) race. */
#define ECS_NOALLOC_SCRATCH_BYTES (64u << 20)
int ecs_state_vel_noalloc(ecs_ctx* ctx, const double* in, double* out, int batch, int include_dissipation,
                          void* stream);

/* Tangent of state_vel: out = v(in), dout = J(in) din  (linearisation of
 * mhd_jax.py:105-175, as jax.jvp of state_vel would give). */
int ecs_state_vel_jvp(ecs_ctx* ctx, const double* in, const double* din, double* out, double* dout,
                      int batch, int include_dissipation, void* stream);

/* Adjoint of state_vel: grad = J(in)^T cot  (what jax.vjp of state_vel gives,
 * with spectral cotangents holding d/dRe in the real part and d/dIm in the
 * imaginary part). */
int ecs_state_vel_vjp(ecs_ctx* ctx, const double* in, const double* cot, double* grad, int batch,
                      int include_dissipation, void* stream);

/* lawson_rk4(f, t, steps, v_fn, L_diag, mask)  (timestepping.py:46-81) with
 * v_fn = state_vel without dissipation and L_diag = dissipation(param_dict),
 * as mismatch_RPO mode "Lawson_RK4" builds them (loss_functions.py:57-65).
 * f: full spectra, updated in place.  dt: host array [batch] of steps h.
 * masked != 0 multiplies e = exp(h/2 L) by the 2/3 mask (the mode's choice). */
int ecs_lawson_rk4(ecs_ctx* ctx, double* f, int batch, const double* dt, int steps, int masked,
                   void* stream);

/* Forward-mode derivative of ecs_lawson_rk4 in (f, h): f and df advance
 * together; ddt: host [batch] tangent of h (dT/steps for h = T/steps). */
int ecs_lawson_rk4_jvp(ecs_ctx* ctx, double* f, double* df, int batch, const double* dt,
                       const double* ddt, int steps, int masked, void* stream);

/* Reverse-mode derivative of ecs_lawson_rk4 (what jax.vjp of the
 * "Lawson_RK4" integration gives, loss_functions.py:57-65).  f0: initial
 * spectra (not modified); fbar_inout: on entry the cotangent of the final
 * spectra, on exit that of f0; hbar: host [batch] output, d/dh.  The forward
 * trajectory is checkpointed every `ckpt` steps and recomputed segment by
 * segment. */
int ecs_lawson_rk4_vjp(ecs_ctx* ctx, const double* f0, double* fbar_inout, double* hbar, int batch,
                       const double* dt, int steps, int masked, int ckpt, void* stream);

/* Two-phase form of ecs_lawson_rk4_vjp for callers that keep the forward
 * result (torch.autograd.Function): ecs_lawson_rk4_fwd_ckpt advances f in
 * place like ecs_lawson_rk4 and stores the compact state at the start of
 * every segment of `ckpt` steps into the caller's device buffer ck of
 * ecs_ckpt_doubles() doubles; ecs_lawson_rk4_vjp_ckpt then runs the reverse
 * pass from ck (recomputing each segment), replacing fbar_inout (cotangent
 * of the final spectra) by that of the initial spectra, and hbar = d/dh.
 * nsteps (host [batch], or NULL): per-candidate step counts, each in
 * [1, steps] and a multiple of ckpt or equal to steps; candidate b stops
 * after nsteps[b] steps of its own dt[b] (serving batches gradient jobs of
 * different periods, adjoint_descent.py:93-123 with steps from T).  Without
 * it every candidate takes `steps`. */
long long ecs_ckpt_doubles(const ecs_ctx* ctx, int batch, int steps, int ckpt);
int ecs_lawson_rk4_fwd_ckpt(ecs_ctx* ctx, double* f, int batch, const double* dt, int steps, int masked, int ckpt,
                            double* ck, const int* nsteps, void* stream);
int ecs_lawson_rk4_vjp_ckpt(ecs_ctx* ctx, const double* ck, double* fbar_inout, double* hbar, int batch,
                            const double* dt, int steps, int masked, int ckpt, const int* nsteps, void* stream);

/* lawson_rk43(x, vel_fn, diag_L, t, h, atol, max_steps_per_checkpoint,
 * checkpoints)  (timestepping.py:151-335).  x: full spectra, updated in
 * place.  T: host [batch].  Synchronous: returns when every candidate has
 * completed or used its budget of checkpoints * max_steps attempts.
 * info: host [batch]; hs: host [batch][budget] steps of each attempt (the
 * reference's info["hs"]); acc: host [batch][budget] 1 if that attempt was
 * accepted (may be NULL). */
int ecs_lawson_rk43(ecs_ctx* ctx, double* x, int batch, const double* T, double h0, double atol,
                    int max_steps_per_checkpoint, int checkpoints, ecs_rk43_info* info, double* hs,
                    int8_t* acc, void* stream);

/* Forward-mode derivative of lawson_rk43 in (x, T) along the schedule a
 * previous ecs_lawson_rk43 call recorded (h is stop-gradient,
 * timestepping.py:308, and decisions follow the primal).  x, dx: full
 * spectra updated in place; T, dT: host [batch]; hs, acc: host
 * [batch][budget] as returned; attempts: host [batch]. */
int ecs_lawson_rk43_jvp(ecs_ctx* ctx, double* x, double* dx, int batch, const double* T,
                        const double* dT, const double* hs, const int8_t* acc, const int* attempts,
                        int budget, void* stream);

/* Reverse-mode derivative of lawson_rk43 in (x, T) along the schedule a
 * previous ecs_lawson_rk43 call recorded (what jax.vjp of mode "Lawson_RK43"
 * gives: h is stop-gradient, timestepping.py:308).  x0: initial spectra (not
 * modified); xbar_inout: cotangent of the final spectra on entry, of x0 on
 * exit; Tbar: host [batch] output d/dT; hs, acc, attempts, budget as for
 * ecs_lawson_rk43_jvp.  Accepted steps are checkpointed every ckpt. */
int ecs_lawson_rk43_vjp(ecs_ctx* ctx, const double* x0, double* xbar_inout, double* Tbar, int batch,
                        const double* T, const double* hs, const int8_t* acc, const int* attempts, int budget,
                        int ckpt, void* stream);

/* The other fixed-step integrators of mismatch_RPO, with their forward-mode
 * derivatives (f, df full spectra updated in place; dt, ddt host [batch], as
 * for ecs_lawson_rk4).  They run as RHS passes with separate stage
 * combinations rather than the fused Lawson-RK4 glue.
 * ecs_rk4:        classical RK4 on the RHS with dissipation
 *                 (timestepping.py:17-44; mode "RK4", loss_functions.py:49-56).
 * ecs_lawson_rk6: Lawson RK6 with e = exp(h/6 L) [* mask], RHS without
 *                 dissipation (timestepping.py:83-125; mode "Lawson_RK6",
 *                 loss_functions.py:66-71; guess generation,
 *                 generate_turbulence.py:88,105).
 * ecs_tdrk4:      two-derivative RK4 with a_j = h J(y_j) k_j, RHS with
 *                 dissipation (timestepping.py:127-149; mode "TDRK4",
 *                 loss_functions.py:78-84; default of newton.py:70).  Its
 *                 tangent carries the second derivative of the RHS. */
int ecs_rk4(ecs_ctx* ctx, double* f, int batch, const double* dt, int steps, void* stream);
int ecs_rk4_jvp(ecs_ctx* ctx, double* f, double* df, int batch, const double* dt, const double* ddt,
                int steps, void* stream);
int ecs_lawson_rk6(ecs_ctx* ctx, double* f, int batch, const double* dt, int steps, int masked, void* stream);
int ecs_lawson_rk6_jvp(ecs_ctx* ctx, double* f, double* df, int batch, const double* dt, const double* ddt,
                       int steps, int masked, void* stream);
int ecs_tdrk4(ecs_ctx* ctx, double* f, int batch, const double* dt, int steps, void* stream);
int ecs_tdrk4_jvp(ecs_ctx* ctx, double* f, double* df, int batch, const double* dt, const double* ddt,
                  int steps, void* stream);

/* Reverse-mode derivatives of the three integrators above (what jax.vjp of
 * mismatch_RPO's integration gives in modes "RK4", "Lawson_RK6", "TDRK4"),
 * with the conventions of ecs_lawson_rk4_vjp: f0 not modified, fbar_inout
 * the cotangent of the final spectra on entry and of f0 on exit, hbar host
 * [batch] = d/dh.  The state is checkpointed every ckpt steps; each segment
 * is run forward again from its checkpoint before it is transposed; TDRK4's
 * transpose carries the second derivative of the RHS. */
int ecs_rk4_vjp(ecs_ctx* ctx, const double* f0, double* fbar_inout, double* hbar, int batch, const double* dt,
                int steps, int ckpt, void* stream);
int ecs_lawson_rk6_vjp(ecs_ctx* ctx, const double* f0, double* fbar_inout, double* hbar, int batch,
                       const double* dt, int steps, int masked, int ckpt, void* stream);
int ecs_tdrk4_vjp(ecs_ctx* ctx, const double* f0, double* fbar_inout, double* hbar, int batch, const double* dt,
                  int steps, int ckpt, void* stream);

/* Two-phase form of the reverse passes above and of ecs_lawson_rk43_vjp, for
 * callers that apply several cotangents at one primal: the vjp function
 * jax.vjp returns (newton.py:168-169), applied once per Krylov step of GMRES
 * on the normal equations (use_transpose, newton.py:167-170) or of the
 * hookstep solvers.  The forward pass runs once and stores the compact state
 * at the start of every segment of ckpt steps into the caller's device
 * buffer ck; each reverse pass then starts from ck.  The same bits as the
 * one-call forms.
 * ecs_gen_fwd_ckpt: scheme 0 classical RK4, 1 Lawson RK6 (masked as for
 *   ecs_lawson_rk6), 2 TDRK4; advances f in place to the final spectra (the
 *   bits of ecs_rk4 / ecs_lawson_rk6 / ecs_tdrk4); ck holds
 *   ecs_ckpt_doubles(ctx, batch, steps, ckpt) doubles.
 * ecs_gen_vjp_ckpt: the reverse pass from ck with the forward pass's
 *   arguments; fbar_inout and hbar as for ecs_rk4_vjp.
 * ecs_lawson_rk43_fwd_ckpt: along the schedule of a previous ecs_lawson_rk43
 *   call (x0 not modified); ck holds ecs_ckpt_doubles(ctx, batch, M, ckpt)
 *   doubles with M the most accepted steps of a candidate (at least 1).
 * ecs_lawson_rk43_vjp_ckpt: the reverse pass from ck; xbar_inout and Tbar as
 *   for ecs_lawson_rk43_vjp. */
int ecs_gen_fwd_ckpt(ecs_ctx* ctx, double* f, int batch, const double* dt, int steps, int scheme, int masked,
                     int ckpt, double* ck, void* stream);
int ecs_gen_vjp_ckpt(ecs_ctx* ctx, const double* ck, double* fbar_inout, double* hbar, int batch, const double* dt,
                     int steps, int scheme, int masked, int ckpt, void* stream);
int ecs_lawson_rk43_fwd_ckpt(ecs_ctx* ctx, const double* x0, int batch, const double* T, const double* hs,
                             const int8_t* acc, const int* attempts, int budget, int ckpt, double* ck, void* stream);
int ecs_lawson_rk43_vjp_ckpt(ecs_ctx* ctx, const double* ck, double* xbar_inout, double* Tbar, int batch,
                             const double* T, const double* hs, const int8_t* acc, const int* attempts, int budget,
                             int ckpt, void* stream);

/* GMRES vector kernels (linalg.py:58-97) on flat FP64 vectors of length n.
 * ecs_mgs: one modified-Gram-Schmidt sweep of Arnoldi step k (linalg.py:70-85):
 * for j = 0..k: h[j] = <w, Q[:, j]>, w -= h[j] Q[:, j]; then h[k+1] = |w| and
 * Q[:, k+1] = w / h[k+1].  Q is column-major with leading dimension ldq
 * (Q[:, j] = Q + j*ldq); h: device [k+2].
 * ecs_mgs_reorth: the same sweep followed by `passes` more orthogonalisation
 * passes against Q[:, 0..k] whose dots are discarded, then the norm and
 * normalisation: the Arnoldi step of the hookstep solvers (linalg.py:137-153
 * and 255-268, two passes).  Reductions use one scratch buffer per (device,
 * stream), so calls on different streams may overlap. */
int ecs_mgs(double* Q, int ldq, int n, int k, double* w, double* h, void* stream);
int ecs_mgs_reorth(double* Q, int ldq, int n, int k, double* w, double* h, int passes, void* stream);
/* out[0] = <a, b> */
int ecs_dot(const double* a, const double* b, int n, double* out, void* stream);

/* Recurrence diagram of a turbulent time series (generate_turbulence.py:140-151):
 * dist[i][j] = || |fs_i| - |fs_j| ||_2 with fs complex128 [nsnap][modes]
 * (each snapshot's spectrum flattened) and dist [nsnap][nsnap], exactly
 * symmetric. */
int ecs_recurrence(const double* fs, int nsnap, long long modes, double* dist, void* stream);

/* Measurement hook (no counterpart in the reference, which has no native
 * code): average duration in ms of one launch of a fused RHS kernel over
 * `reps` back-to-back launches on `batch` resident spectra, timed with HIP
 * events on `stream`.  which: 0 column kernel, 1 row kernel, 2 tangent
 * column kernel, 3 tangent row kernel, 4 tangent row kernel of the adaptive
 * tangent replay (the instance of a GMRES matvec), 5 an empty kernel with the
 * tangent row kernel's grid and argument block (the launch floor).
 * Synchronous. */
int ecs_time_kernel(ecs_ctx* ctx, int batch, int which, int reps, double* ms, void* stream);

/* Measurement hook (no counterpart in the reference): with enable = 1, the
 * following ecs_lawson_rk43_jvp calls on ctx record HIP events on their
 * stream around the first tangent row pass of every `every`-th accepted step
 * (up to 1024 samples); enable = 0 stops recording, waits for the last event
 * and returns the average duration in ms of the sampled launches, as they
 * run inside the replay (after a column kernel, caches as the replay leaves
 * them), and the number of samples. */
int ecs_replay_probe(ecs_ctx* ctx, int enable, int every, double* avg_ms, int* samples);

/* Tile choice of the autotuner for grid size n and this batch bucket (no
 * reference counterpart): out[0] columns per column-kernel block, out[1] the
 * same for the tangent column kernel, out[2] rows per row-kernel block,
 * out[3] the same for the tangent row kernel; -1 where not tuned yet. */
int ecs_tuned_tiles(int n, int batch, int* out);

/* FFT radix plan the autotuner fixed for grid size n (no reference
 * counterpart), per kernel family in the order of ecs_tuned_tiles: 16 for
 * the wide plan (radix 16/8: 128 = 16x8, 256 = 16x16), 8 for the narrow one
 * (radix 8/4/2: 128 = 8x4x4, 256 = 8x8x4); -1 where not tuned yet. */
int ecs_tuned_plans(int n, int* out);

/* Column split of the autotuner for grid size n and this batch bucket (no
 * reference counterpart): out[0] for the column kernel, out[1] for the
 * tangent column kernel, the number of blocks that share one column's output
 * sequences (1, 2 or 4; a split variant has one column per block); -1 where
 * not tuned yet. */
int ecs_tuned_splits(int n, int batch, int* out);

/* Raw variant indices the autotuner chose for grid size n and this batch
 * bucket (no reference counterpart), per kernel family in the order of
 * ecs_tuned_tiles; -1 where not tuned yet.  v < 6: tile 1 << (v % 3), wide
 * plan for v < 3, narrow for 3..5; column families 6..9: one column per block
 * split over 2 (6, 8) or 4 (7, 9) blocks, wide plan for 6, 7; for n >= 128
 * also 10 and 11: tiles of 8 columns, wide and narrow plan; 12..15: tiles of
 * 2 columns split over 2 (12, 14) or 4 (13, 15) blocks, wide plan for 12,
 * 13.  The same
 * indices are accepted by ECSMHD_VARIANTS="N:c,ct,r,rt". */
int ecs_tuned_variants(int n, int batch, int* out);

/* Raw variant indices (as ecs_tuned_variants) the autotuner chose for the
 * tangent column (out[0]) and tangent row kernel (out[1]) as the tangent
 * replay of an adaptive schedule runs them (ecs_lawson_rk43_jvp, the GMRES
 * matvec; no reference counterpart): tuned in that instance, apart from the
 * fixed-step tangent kernels; -1 where not tuned yet.  ECSMHD_VARIANTS sets
 * them to its ct and rt. */
int ecs_tuned_replay_variants(int n, int batch, int* out);

#ifdef __cplusplus
}
#endif
#endif
