/*
 * socu_hip.h -- C ABI of the MI355X (gfx950) nested-dissection block-tridiagonal
 * Cholesky library, libsocu_hip.so.
 *
 * The reference (PREDICT-EPFL/socu) is pure Python over NVIDIA Warp; its
 * "native boundary" is the set of launch factories in
 * socu/block_tridiag_solver.py.  Each entry point below replaces the device
 * work one of those factories enqueues, so the Python mirror in
 * socu_amd/block_tridiag_solver.py can keep the reference's signatures.
 *
 * Conventions (all entry points):
 *  - L, E, x are device pointers to contiguous row-major arrays
 *      L (B, N, n, n), E (B, n_E, n, n) with n_E = socu_calculate_off_diag_storage_len(N),
 *      x (B, N, n, m);
 *    dtype selects the element type (SOCU_DTYPE_F32 = float, SOCU_DTYPE_F64 = double).
 *  - mask is a device pointer to B bytes or NULL; a zero byte leaves that
 *    batch entry's L, E and x untouched (README.md:214-235).  It is read by
 *    the kernels on every execution, including graph replays.
 *  - stream is a hipStream_t (NULL = the null stream).  Work is enqueued
 *    asynchronously; nothing synchronises the host.  Only the *_create
 *    calls allocate device memory (plan workspaces, graphs); the direct
 *    entry points and every *_execute allocate nothing.
 *  - Return 0 on success, a negative SOCU_E* code otherwise;
 *    socu_last_error() returns the calling thread's last message.
 */
#ifndef SOCU_HIP_H
#define SOCU_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SOCU_DTYPE_F32 0
#define SOCU_DTYPE_F64 1

/* Factor precisions of FP32 storage (socu_cholesky_factor_f32, the
 * mixed-precision solver); see there. */
#define SOCU_PREC_FP32 0
#define SOCU_PREC_FP32_F32PIV 1
#define SOCU_PREC_BF16 2

#define SOCU_OK 0
#define SOCU_EINVAL -1      /* bad argument (shape, dtype, pointer) */
#define SOCU_EUNSUPPORTED -2 /* valid but not supported (e.g. n > max block size) */
#define SOCU_EHIP -3        /* a HIP runtime call failed */
#define SOCU_EEXPIRED -4    /* a bounded wait expired in an earlier execution of this plan / handle */

/* Operations a plan can hold (the five launch factories). */
#define SOCU_OP_FACTOR 0           /* create_cholesky_factor_launch, block_tridiag_solver.py:739-923 */
#define SOCU_OP_SOLVE 1            /* create_cholesky_solve_launch, :1520-1727 */
#define SOCU_OP_FACTOR_AND_SOLVE 2 /* create_cholesky_factor_and_solve_launch, :926-1233 */
#define SOCU_OP_FORWARD 3          /* create_cholesky_forward_substitution_launch, :1236-1374 */
#define SOCU_OP_BACKWARD 4         /* create_cholesky_backward_substitution_launch, :1377-1517 */

/* calculate_recursive_iterations, block_tridiag_solver.py:12-15 (N >= 1). */
int socu_calculate_recursive_iterations(int N);

/* calculate_off_diag_storage_len, block_tridiag_solver.py:18-23 (N >= 1). */
long long socu_calculate_off_diag_storage_len(int N);

/* Largest block size n the kernels accept. */
int socu_max_block_size(void);

/* In-place factor of every unmasked system: tril(L[b,i]) <- diagonal Cholesky
 * blocks, E[b] <- ND off-diagonal factor blocks including fill-in.
 * Replaces the device work of create_cholesky_factor_launch's callback
 * (block_tridiag_solver.py:864-918).  The reference needs the fill-in region
 * E[b, N-1:] zeroed before each factor (it accumulates F -= R Lc into it);
 * here each fill-in block is written once as F = -R Lc, so its contents on
 * entry are ignored (zeros give the same result). */
int socu_cholesky_factor(void* L, void* E, const unsigned char* mask, int B, int N, int n, int dtype, void* stream);

/* socu_cholesky_factor in FP32 storage with the precision options of the
 * mixed-precision solver (SOCU_PREC_*, below): SOCU_PREC_FP32 factors and
 * inverts the diagonal blocks in FP64 (rounded to FP32 on store), SOCU_PREC_BF16
 * does so too and runs the Schur-complement products (syrk and fill-in gemm)
 * on BF16 MFMA with FP32 accumulation, SOCU_PREC_FP32_F32PIV is the plain
 * FP32 factor.  The reference's factor offers FP32 or FP64 end to end
 * (block_tridiag_solver.py:26-39); this is the factor the mixed solver
 * refines from, for callers that keep their own refinement. */
int socu_cholesky_factor_f32(void* L, void* E, const unsigned char* mask, int B, int N, int n, int precision,
                             void* stream);

/* Substitutions read only tril(L) and E (as the reference's sweeps do,
 * socu/blas/potrf.py:101-103), so the factor may come from anywhere.  They
 * first form the inverses of the diagonal factor blocks from tril(L) into
 * `work`, a device buffer of socu_solve_workspace_size(B, N, n, dtype) bytes
 * (one n x n block per diagonal block), which they own for the duration of
 * the call; L and E are not modified.  work must not be NULL when B > 0. */
size_t socu_solve_workspace_size(int B, int N, int n, int dtype);

/* x <- Lp^{-1} P x (forward sweep), block_tridiag_solver.py:1339-1369. */
int socu_cholesky_forward_substitution(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N,
                                       int n, int m, int dtype, void* work, void* stream);

/* x <- P^T Lp^{-T} x (backward sweep), block_tridiag_solver.py:1480-1512. */
int socu_cholesky_backward_substitution(const void* L, const void* E, void* x, const unsigned char* mask, int B,
                                        int N, int n, int m, int dtype, void* work, void* stream);

/* forward then backward sweep, block_tridiag_solver.py:1666-1722. */
int socu_cholesky_solve(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N, int n, int m,
                        int dtype, void* work, void* stream);

/* forward then backward sweep with a factor made by this library's
 * socu_cholesky_factor or socu_cholesky_factor_and_solve, which leave the
 * transposed inverse of each diagonal factor block in the strict upper
 * triangle of L[b, i]: the sweeps read it there instead of forming it from
 * tril(L), so no workspace is needed and the inverse launch is saved (no
 * counterpart in the reference, whose L upper triangle is unspecified; a
 * factor made any other way needs socu_cholesky_solve).  For callers that
 * keep the factor they made, e.g. the backward pass of a differentiable
 * solve. */
int socu_cholesky_solve_with_inverses(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N,
                                      int n, int m, int dtype, void* stream);

/* socu_cholesky_solve without a workspace argument and without any device
 * allocation: no counterpart in the reference; it is for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same kernels and results as socu_cholesky_solve; the
 * diagonal-block inverses go to a static device array of
 * SOCU_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends
 * (its address is looked up on the first call on a device).  Returns
 * SOCU_EUNSUPPORTED when socu_solve_workspace_size(B, N, n, dtype) exceeds
 * it (64 MiB: e.g. B*N <= 512 FP64 blocks of n = 128).  Every call on one
 * GPU uses the same array, so calls that overlap on one GPU (other streams,
 * other threads) race; calls in order on one stream are safe. */
#define SOCU_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int socu_cholesky_solve_noalloc(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N, int n,
                                int m, int dtype, void* stream);

/* factor, forward, backward, block_tridiag_solver.py:1124-1228.  The
 * factor leaves the transposed inverse of each diagonal factor block in the
 * strict upper triangle of L[b, i] (the reference leaves that triangle
 * implementation-defined), and the sweeps of the same call read it there. */
int socu_cholesky_factor_and_solve(void* L, void* E, void* x, const unsigned char* mask, int B, int N, int n, int m,
                                   int dtype, void* stream);

/* Out-of-place factor and factor_and_solve.  No counterpart in the
 * reference (whose factor is in place, block_tridiag_solver.py:864-918);
 * for callers that must keep their inputs, such as a differentiable solve
 * (the backward pass needs the factor, the caller's L and E stay as given):
 * the factor of (L_src, E_src) is written to (L, E), which need not hold
 * anything on entry, without copying the inputs first (the first level reads
 * them, every later one reads L and E).  Only tril(L_src) and the coupling
 * blocks E_src[b, :N-1] are read; the sources are not written.  L / E must
 * not overlap L_src / E_src.  Masked-out systems are not written.  Results
 * equal those of socu_cholesky_factor / socu_cholesky_factor_and_solve on
 * copies of the inputs.  (For n > 128 the inputs are copied, then factored
 * in place.) */
int socu_cholesky_factor_from(const void* L_src, const void* E_src, void* L, void* E, const unsigned char* mask, int B,
                              int N, int n, int dtype, void* stream);
int socu_cholesky_factor_and_solve_from(const void* L_src, const void* E_src, void* L, void* E, void* x,
                                        const unsigned char* mask, int B, int N, int n, int m, int dtype,
                                        void* stream);

/* A plan binds one operation to its arrays, like the reference's launch
 * callables (arrays bound at creation).  With use_graph != 0 the work is
 * captured once into a hipGraph on a private capture stream and replayed by
 * socu_plan_execute (replaces utils.py:40-66, create_cuda_graph_callback).
 * x may be NULL for SOCU_OP_FACTOR.  Solve, forward and backward plans own
 * the diagonal-block inverse workspace of the direct entry points. */
typedef struct socu_plan socu_plan;
int socu_plan_create(socu_plan** out, int op, void* L, void* E, void* x, const unsigned char* mask, int B, int N,
                     int n, int m, int dtype, int use_graph);
int socu_plan_execute(socu_plan* plan, void* stream);
int socu_plan_destroy(socu_plan* plan);

/* Substitution sweeps with one right-hand side (m == 1): a plan (and a mixed
 * handle) owns a small workspace and runs the upper ND levels of forward and
 * backward substitution as one persistent launch of node tasks that wait on
 * per-node counters; the wide lower levels keep the level kernels (same
 * arithmetic per block as the reference's sweeps,
 * block_tridiag_solver.py:425-485 and :581-643).  Used by every plan with one
 * right-hand side, in both precisions (it measured faster than the level
 * kernels for every block size); the environment variable SOCU_SWEEP read at
 * creation overrides: "level" (never) or "persistent" (always).  Each wait
 * is bounded; if one expires the launch drains, the tasks that depended on
 * the wait write NaN into their part of x (so the failure shows in the
 * solution), and a host-visible word is raised: the plan's next
 * socu_plan_execute then returns SOCU_EEXPIRED without running (and clears
 * the word).  socu_plan_sweep_status synchronises `stream` and reports
 * whether the plan uses the persistent sweeps (*persistent) and whether a
 * wait expired in its last execution (*expired; this also clears the word).
 * Executions of one plan must not overlap (they share the workspace).  (No
 * counterpart in the reference.) */
int socu_plan_sweep_status(socu_plan* plan, int* persistent, int* expired, void* stream);
/* Polls per dependency wait of the persistent sweeps enqueued from now on
 * (process-wide; a captured graph keeps the value it was captured with).
 * 0 restores the default (2^21 polls); a negative value makes every wait that
 * is not already satisfied expire at once.  For tests of the failure path:
 * every kernel still exits normally.  (No counterpart in the reference.) */
int socu_set_sweep_spin_limit(int polls);

/* Mixed-precision factor_and_solve with FP64 iterative refinement (no
 * counterpart in the reference; an extension with the same data layout).
 * L, E (FP64, the matrix Psi) are left untouched; x holds b on entry and the
 * FP64 solution on exit.  The handle owns an FP32 working copy of (L, E) that
 * is factored with FP32 MFMA Schur updates (precision SOCU_PREC_FP32: the
 * diagonal blocks are factored and inverted in FP64, then rounded;
 * SOCU_PREC_FP32_F32PIV: all in FP32; SOCU_PREC_BF16: FP64 diagonal blocks as
 * SOCU_PREC_FP32, the Schur-complement products (syrk and fill-in gemm) on
 * BF16 MFMA with FP32 accumulation), then runs up to max_iter correction
 * solves d = Lp32^-T Lp32^-1 r, x += d, r = b - Psi x (FP64), stopping per
 * system once ||r||_F <= tol ||b||_F.  The test runs on the device, so the
 * whole solve is one hipGraph (use_graph) with no host synchronisation.
 * socu_mixed_create allocates the workspace; execute allocates nothing.
 * socu_mixed_stats synchronises `stream` and returns, per system, the last
 * relative residual and the number of correction solves applied. */
typedef struct socu_mixed socu_mixed;
int socu_mixed_create(socu_mixed** out, const void* L, const void* E, void* x, const unsigned char* mask, int B, int N,
                      int n, int m, int precision, double tol, int max_iter, int use_graph);
int socu_mixed_execute(socu_mixed* handle, void* stream);
int socu_mixed_stats(socu_mixed* handle, double* rel_resid, int* solves, void* stream);
int socu_mixed_destroy(socu_mixed* handle);
/* As socu_plan_sweep_status, for the FP32 correction solves of a handle
 * (socu_mixed_execute likewise returns SOCU_EEXPIRED after an execution whose
 * wait expired). */
int socu_mixed_sweep_status(socu_mixed* handle, int* persistent, int* expired, void* stream);
/* Runs the mixed solve once without the graph, with the per-kernel-class
 * event timer of socu_profile_op (classes 0-4 as there, 5 = FP64 residual and
 * refinement updates), and synchronises `stream`. */
int socu_mixed_profile(socu_mixed* handle, void* stream, double* kernel_ms, int* kernel_launches);

/* Instrumentation (no counterpart in the reference): runs `op` once like the
 * direct entry points, with a hipEvent pair around every kernel launch on
 * `stream`, synchronises `stream`, and returns per kernel class the summed
 * elapsed milliseconds and the launch count (arrays of SOCU_KERNEL_CLASSES).
 * Classes: 0 potrf, 1 trsm, 2 Schur update, 3 forward step, 4 backward level,
 * 5 residual / refinement.  Creates events, so it is
 * for measurement, not for hot loops. */
#define SOCU_KERNEL_CLASSES 6
int socu_profile_op(int op, void* L, void* E, void* x, const unsigned char* mask, int B, int N, int n, int m,
                    int dtype, void* work, void* stream, double* kernel_ms, int* kernel_launches);

/* Batched uneven systems (no counterpart in the reference, whose launches
 * batch only systems sharing N and n: block_tridiag_solver.py:749-768).
 * Each entry is one independent system in the reference's layout: L (N,n,n),
 * E (calculate_off_diag_storage_len(N),n,n) (the fill-in region's contents on
 * entry are ignored, as in socu_cholesky_factor), and x
 * (N,n,m) (ignored for SOCU_OP_FACTOR); `reserved` must be 0.  Needs
 * n <= socu_batched_max_block_size() and m <= 32.  mask is NULL or `count`
 * bytes indexed like `systems`.  One workgroup per system walks all its
 * levels.  Systems fall in size buckets by max(n, m) (<= 8, <= 16, <= 32);
 * each bucket is one launch, with a tile size of at least its bound and
 * workgroups of 1, 2 or 4 wavefronts, the launches on forked streams.  With
 * `tune` the create call times every admissible (tile size, wavefronts) per
 * bucket on a device copy of the batch (the caller's arrays are only read)
 * and keeps the fastest (times within 2% of it are ties, settled for the
 * more wavefronts); results are cached per process and, if the
 * environment variable SOCU_TUNE_CACHE names a file, appended there and
 * reused by later processes.  Without `tune` each bucket uses its own bound
 * and two wavefronts (or the count in the environment variable
 * SOCU_BATCHED_WAVES: 1, 2 or 4), fewer where the LDS does not fit.  create
 * copies the system table to the device and
 * synchronises; execute allocates nothing and synchronises nothing.
 * socu_batched_config returns per bucket the system count, tile size and
 * wavefronts per workgroup. */
typedef struct socu_batch_system {
  void* L;
  void* E;
  void* x;
  int N;
  int n;
  int m;
  int reserved;
} socu_batch_system;
typedef struct socu_batched socu_batched;
int socu_batched_max_block_size(void);
int socu_batched_create(socu_batched** out, int op, const socu_batch_system* systems, int count,
                        const unsigned char* mask, int dtype, int tune, int use_graph);
int socu_batched_execute(socu_batched* handle, void* stream);
int socu_batched_config(socu_batched* handle, int* counts, int* tiles, int* waves);
int socu_batched_destroy(socu_batched* handle);

/* Adjoint scatter (no counterpart in the reference, which has no gradient).
 * For x = Psi^{-1} b, with Psi the block-tridiagonal matrix of the solver
 * (diagonal blocks from the lower triangle of L_i, couplings E_i at (i+1, i))
 * and bbar = Psi^{-1} xbar (one solve with the same factor), writes the
 * gradient of the loss with respect to the parameters:
 *   dL_i = lower part of -(bbar_i x_i^T + x_i bbar_i^T), halved on the
 *          diagonal, zero above it;
 *   dE_e = -(bbar_{e+1} x_e^T + x_{e+1} bbar_e^T) for e < N-1, zero on the
 *          fill-in region.
 * Shapes as the solver: bbar, x (B,N,n,m); dL (B,N,n,n); dE (B,n_E,n,n).
 * Masked-out systems keep their dL, dE. */
int socu_adjoint_scatter(const void* bbar, const void* x, void* dL, void* dE, const unsigned char* mask, int B, int N,
                         int n, int m, int dtype, void* stream);

/* Gradient with respect to the entries of a stored ND factor (no counterpart
 * in the reference).  With Lp the ND-permuted Cholesky factor held in
 * (tril(L), E) as the factor leaves it, writes G = -(u v^T + p q^T)
 * restricted to the entries of Lp, in the factor's own storage:
 *   dL_i = tril(G_ii) (diagonal included, strict upper triangle zero);
 *   E entry k of level segment l (coupling a = (k+1)s-1, b = (k+2)s-1,
 *   s = 2^l, block_tridiag_solver.py:92-145) holds Lp[b][a] for even k and
 *   Lp[a][b]^T for odd k, so dE = -(u_b v_a^T + p_b q_a^T) for even k and
 *   -(v_b u_a^T + q_b p_a^T) for odd k.
 * u, v, p, q are (B,N,n,m) in the solver's vector layout; p and q may both be
 * NULL.  For y = forward_substitution(b): u = backward_substitution(ybar),
 * v = y.  For x = backward_substitution(y): u = x, v = forward_substitution(xbar).
 * A solve is the sum of both.  Masked-out systems keep their dL, dE. */
int socu_factor_adjoint_scatter(const void* u, const void* v, const void* p, const void* q, void* dL, void* dE,
                                const unsigned char* mask, int B, int N, int n, int m, int dtype, void* stream);

/* Last error message of the calling thread ("" if none). */
const char* socu_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SOCU_HIP_H */
