// Host-side launchers shared between the kernel files and the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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

// One-time per-process kernel set-up (LDS limits above 64 KiB).  Must run
// before any stream capture.
hipError_t init_blocked_kernels();

// Kernels of one ND factor level (potrf, then trsm and Schur updates unless it
// is the last level), for block size n <= kMaxN.
// f64_pivots (float storage only): the diagonal blocks are factored and
// inverted in FP64 and rounded to FP32 on store.
template <typename T>
hipError_t factor_level_blocked(T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, bool f64_pivots,
                                hipStream_t st);

// Forward-substitution step: step -1 applies level 0; step l >= 0 applies
// level l's updates to every receiving block, then level l+1's inverses.
template <typename T>
hipError_t forward_step_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                int step, hipStream_t st);

// Backward-substitution kernels of one ND level.
template <typename T>
hipError_t backward_level_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                  int lidx, hipStream_t st);

// Mixed-precision factor_and_solve with FP64 refinement (nd_refine.hip).
struct MixedPlan {
  const double* L = nullptr;  // Psi, FP64, untouched
  const double* E = nullptr;
  double* x = nullptr;        // b on entry, solution on exit
  const uint8_t* mask = nullptr;
  int B = 0, N = 0, n = 0, m = 0, nE = 0;
  double tol = 1e-10;
  int max_iter = 10;
  bool f64_pivots = true;
  // workspace
  float* L32 = nullptr;
  float* E32 = nullptr;
  float* d32 = nullptr;
  double* b64 = nullptr;
  double* state = nullptr;    // 4 per system: ||b||^2, ||r||^2, done, solves
  double* partial = nullptr;
  uint8_t* active = nullptr;
  int nblk_init = 1;
};
hipError_t init_refine_kernels();
size_t mixed_residual_lds(int n, int m);
hipError_t mixed_alloc(MixedPlan& p);
void mixed_free(MixedPlan& p);
hipError_t mixed_enqueue(const MixedPlan& p, hipStream_t st);

}  // namespace socu
