// Host-side launchers shared between the kernel files and the C-ABI.
#pragma once

#include <functional>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nd_common.h"

namespace socu {

// Kernel classes for the optional per-launch event timer (socu_profile_op).
enum KernelClass {
  kKcPotrf = 0,
  kKcTrsm = 1,
  kKcSchur = 2,
  kKcForward = 3,
  kKcBackward = 4,
  kKcRefine = 5,
  kKcCount = 6
};

// Record a timing event before / after a launch of class `cls` on `st`.
// No-ops unless socu_profile_op is running on this thread.
void kt_begin(int cls, hipStream_t st);
void kt_end(int cls, hipStream_t st);
// Under socu_profile_op: a begin/end event pair of class `cls` for a launch
// that records them itself (hipExtLaunchKernel), and true; else false.
bool kt_kernel_events(int cls, hipEvent_t* start, hipEvent_t* stop);

// One-time per-process kernel set-up (LDS limits above 64 KiB).  Must run
// before any stream capture.
hipError_t init_blocked_kernels();

// Precision options of a factor in float storage (bit flags; ignored for
// double storage):
//   kPrecF64Pivots: the diagonal blocks are factored and inverted in FP64 and
//     rounded to FP32 on store;
//   kPrecBf16Schur: the Schur-complement products (syrk and the fill-in gemm)
//     run on BF16 MFMA with FP32 accumulation, operands rounded from FP32.
constexpr int kPrecF64Pivots = 1;
constexpr int kPrecBf16Schur = 2;

// Kernels of one ND factor level (potrf, then trsm and Schur updates unless it
// is the last level), for block size n <= kMaxNLarge; prec: kPrec* flags
// (n <= kMaxN only).
// after_potrf / after_trsm (optional) are recorded on st after the level's
// potrf and trsm launches (the forward sweep can follow them on another
// stream); st waits for before_trsm (optional) between potrf and trsm.
// Lsrc / Esrc (level 0 only, n <= kMaxN, both or neither): the level reads
// the input blocks from them instead of L / E and writes its results to
// L / E -- every L block and every coupling block of E is read first by
// level 0, so this makes the whole factor out of place.
template <typename T>
hipError_t factor_level_blocked(T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, int prec,
                                hipStream_t st, hipEvent_t after_potrf = nullptr, hipEvent_t after_trsm = nullptr,
                                hipEvent_t before_trsm = nullptr, const T* Lsrc = nullptr, const T* Esrc = nullptr);

// Per-node kernels for kMaxN < n <= kMaxNLarge (nd_large.hip): potrf with the
// transposed inverse in the strict upper triangle; trsm of level lidx's R and
// Lc; the forward (U^T x) or backward (U x) apply with U = Linv^T as stored,
// for level lidx's nodes; the sweep updates of level lidx; the inverse of a
// factor from tril(L) alone into W.
hipError_t init_large_kernels();
template <typename T>
hipError_t potrf_large(T* L, const uint8_t* mask, int B, int N, int n, int lidx, hipStream_t st);
template <typename T>
hipError_t trsm_large(const T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, hipStream_t st);
template <typename T>
hipError_t apply_large(const T* L, T* x, const uint8_t* mask, int B, int N, int n, int m, int lidx, bool bwd,
                       hipStream_t st);
template <typename T>
hipError_t update_large(const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, int lidx, bool bwd,
                        hipStream_t st);
template <typename T>
hipError_t trtri_large(const T* L, T* W, const uint8_t* mask, int B, int N, int n, hipStream_t st);

// W <- tril(L) with the transposed inverse of each diagonal block in its
// strict upper triangle (the layout the factor leaves in L), from tril(L)
// only; one launch over the B*N blocks.  Solves of a factor that need not
// come from this library's factor read W in place of L.
template <typename T>
hipError_t trtri_blocked(const T* L, T* W, const uint8_t* mask, int B, int N, int n, hipStream_t st);

// Forward-substitution step: step -1 applies level 0; step l >= 0 applies
// level l's updates to every receiving block, then (unless apply_next is
// false) level l+1's inverses.
template <typename T>
hipError_t forward_step_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                int step, hipStream_t st, bool apply_next = true);

// The two halves of a forward step: level `step`'s updates into every
// receiving block (needs that level's trsm) and the inverses of level
// `level` (needs its potrf and every update into it).
template <typename T>
hipError_t forward_update_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                  int step, hipStream_t st);
template <typename T>
hipError_t forward_apply_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                 int level, hipStream_t st);

// Backward-substitution kernels of one ND level.
template <typename T>
hipError_t backward_level_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                  int lidx, hipStream_t st);

// Persistent dataflow sweeps for one right-hand side (nd_sweep.hip).
// Workspace: sweep_ws_ints(B, N) ints (ticket, status word, per-node
// counters; cleared by the launch itself) and sweep_slot_elems(B, N, n)
// elements of the solve's type (one n-vector per contribution).
constexpr int kSweepHeader = 4;             // ticket, status, 2 spare
constexpr int kSweepSpinLimit = 1 << 21;    // polls per dependency wait
constexpr int kSweepMaxLevels = 32;         // levels of any N < 2^31
// Contribution slots of one system: two per receiving block and step,
// sum over levels l of 2 * (N >> (l+1)).
__host__ __device__ inline int sweep_slots_per_system(int N) {
  int t = 0;
  for (int l = 0; (N >> l) > 0; ++l) t += 2 * (N >> (l + 1));
  return t;
}
hipError_t init_sweep_kernels();
size_t sweep_ws_ints(int B, int N);
size_t sweep_slot_elems(int B, int N, int n);
// Forward and/or backward substitution of x (B, N, n, 1) with the factor in
// L, E; one clear launch (skipped with clear = false, for callers that zero
// the workspace themselves) and one persistent launch.
// hstatus (optional): a host-visible word (pinned host memory) that an
// expired wait sets to 1, so the host can see it without synchronising.
template <typename T>
hipError_t sweep_persistent(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, bool fwd, bool bwd,
                            int* ws, T* slots, int* hstatus, hipStream_t st, bool clear = true);
// Polls per dependency wait for kernels enqueued from now on; 0 restores the
// default (kSweepSpinLimit), a negative value makes every wait that is not
// already satisfied expire at once (for tests of the failure path).
void set_sweep_spin_limit(int polls);
int sweep_spin_limit();
// Whether a plan for block size n and element size dsize uses the persistent
// sweeps (plans with one right-hand side only): yes, they measured faster than
// the level kernels for every block size and precision, unless the
// environment variable SOCU_SWEEP is "level" (never) or "persistent"
// (always, also the default).
bool sweep_persistent_enabled(int n, int dsize);

// A side stream and per-level events that let the forward sweep of a
// factor_and_solve run alongside the factor: forward step l waits for level
// l's trsm, the inverses of level l for its potrf (created with a plan,
// destroyed with it).
struct SweepOverlap {
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, ready = nullptr;
  hipEvent_t potrf[kSweepMaxLevels] = {};
  hipEvent_t trsm[kSweepMaxLevels] = {};
};
hipError_t overlap_create(SweepOverlap& ov);
void overlap_destroy(SweepOverlap& ov);
// Factor levels on st with the forward sweep of x on ov.side, joined back
// into st at the end (the backward sweep is the caller's).  side_first
// (optional) is enqueued on ov.side before the sweep; the first level's trsm
// waits for it (it may prepare E, and anything the sweep reads).
template <typename T>
hipError_t factor_with_forward(T* L, T* E, T* x, const uint8_t* factor_mask, const uint8_t* sweep_mask, int B, int N,
                               int n, int m, int prec, const SweepOverlap& ov, hipStream_t st,
                               const std::function<hipError_t(hipStream_t)>* side_first = nullptr,
                               const T* Lsrc = nullptr, const T* Esrc = nullptr);

// n == 32 (blocks in the storage precision): the factor levels with the
// forward sweep of x inside the same kernels (k_node32, k_schur_small), one
// stream, no separate sweep kernels; x then holds y for the backward sweep.
// fused_forward_applies says whether (n, prec) takes this path;
// factor_and_forward returns hipErrorInvalidValue otherwise.  Lsrc / Esrc as
// for factor_level_blocked.
template <typename T>
bool fused_forward_applies(int n, int prec);
template <typename T>
hipError_t factor_and_forward(T* L, T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, int prec,
                              hipStream_t st, const T* Lsrc = nullptr, const T* Esrc = nullptr);

// Mixed-precision factor_and_solve with FP64 refinement (nd_refine.hip).
struct MixedPlan {
  const double* L = nullptr;  // Psi, FP64, untouched
  const double* E = nullptr;
  double* x = nullptr;        // b on entry, solution on exit
  const uint8_t* mask = nullptr;
  int B = 0, N = 0, n = 0, m = 0, nE = 0;
  double tol = 1e-10;
  int max_iter = 10;
  int prec = kPrecF64Pivots;  // kPrec* flags of the FP32 factor
  // workspace
  float* L32 = nullptr;
  float* E32 = nullptr;
  float* d32 = nullptr;
  double* b64 = nullptr;
  double* state = nullptr;    // 4 per system: ||b||^2, ||r||^2, done, solves
  double* partial = nullptr;
  uint8_t* active = nullptr;
  int nblk_init = 1;
  int* sweep_ws = nullptr;     // persistent sweeps (m == 1), else NULL
  float* sweep_slots = nullptr;
  int* sweep_hstatus = nullptr;  // pinned host word raised by an expired wait
  int* sweep_hstatus_dev = nullptr;  // its device address
  double* res_w = nullptr;     // one-RHS residual pieces (m == 1), else NULL
  double* res_v = nullptr;
  double* res_c = nullptr;
  int* res_ticket = nullptr;   // per system: residual workgroups done
  bool use_overlap = false;     // set before mixed_alloc
  SweepOverlap overlap;        // first correction solve's forward sweep alongside the factor
};
hipError_t init_refine_kernels();

// Batched systems of uneven size (nd_batched.hip): one entry per system, all
// arrays contiguous row-major in the system's own N and n (n <= 32, m <= 32).
struct BatchProblem {
  void* L;  // (N, n, n)
  void* E;  // (n_E(N), n, n)
  void* x;  // (N, n, m), may be NULL for a factor
  int N, n, m, pad;
};
hipError_t init_batched_kernels();
// Tile class of a block size: 0 (tiles of 8), 1 (16), 2 (32).
int batched_class(int n);
// One launch over `count` systems with tile class `cls` and workgroups of
// `waves` (1, 2 or 4) wavefronts, listed (by index into dprobs) at dorder;
// mmax is the largest m among them (<= 8 << cls).  Fails when that
// workgroup's LDS would exceed a compute unit's (batched_fits).
bool batched_fits(int cls, int waves, int mmax, int dtype);
hipError_t batched_launch_class(const BatchProblem* dprobs, const int* dorder, int cls, int count, int mmax, int waves,
                                const uint8_t* mask, int op, int dtype, hipStream_t st);
// Copies L, E (and x unless op is a factor) of count systems src[k] -> dst[k].
hipError_t batched_copy(const BatchProblem* dsrc, const BatchProblem* ddst, int count, int op, int dtype,
                        hipStream_t st);
size_t mixed_residual_lds(int n, int m);
bool mixed_fits(int n, int m);
hipError_t mixed_alloc(MixedPlan& p);
void mixed_free(MixedPlan& p);
hipError_t mixed_enqueue(const MixedPlan& p, hipStream_t st);

// Adjoint scatter (nd_adjoint.hip): dL, dE of x = Psi^{-1} b from
// bbar = Psi^{-1} xbar and x; all arrays in the solver's (B, ...) layout.
hipError_t adjoint_scatter(const void* bbar, const void* x, void* dL, void* dE, const uint8_t* mask, int B, int N,
                           int n, int m, int dtype, hipStream_t st);
// Gradient with respect to the stored ND factor entries:
// G = -(u v^T + p q^T) on the pattern of Lp (p, q may both be NULL).
hipError_t factor_adjoint_scatter(const void* u, const void* v, const void* p, const void* q, void* dL, void* dE,
                                  const uint8_t* mask, int B, int N, int n, int m, int dtype, hipStream_t st);

}  // namespace socu
