// Host-side launchers shared between the kernel files and the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace socu {

// Kernel classes for the optional per-launch event timer (socu_profile_op).
enum KernelClass { kKcPotrf = 0, kKcTrsm = 1, kKcSchur = 2, kKcForward = 3, kKcBackward = 4, kKcCount = 5 };

// Record a timing event before / after a launch of class `cls` on `st`.
// No-ops unless socu_profile_op is running on this thread.
void kt_begin(int cls, hipStream_t st);
void kt_end(int cls, hipStream_t st);

// One-time per-process kernel set-up (LDS limits above 64 KiB).  Must run
// before any stream capture.
hipError_t init_factor_kernels();
hipError_t init_solve_kernels();
hipError_t init_blocked_kernels();

// Blocked kernels for n <= 128 (nd_blocked.hip): same contracts as the
// launchers below.
template <typename T>
hipError_t factor_level_blocked(T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, hipStream_t st);
template <typename T>
hipError_t forward_step_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                int step, hipStream_t st);
template <typename T>
hipError_t backward_level_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                  int lidx, hipStream_t st);

// Factor kernels of one ND level (potrf, then trsm and Schur updates unless
// it is the last level).
template <typename T>
hipError_t factor_level(T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, hipStream_t st);

// Forward-substitution step: step -1 solves level 0; step l >= 0 applies
// level l's updates to every receiving block and solves level l+1.
template <typename T>
hipError_t forward_step(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, int step,
                        hipStream_t st);

// Backward-substitution kernel of one ND level.
template <typename T>
hipError_t backward_level(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, int lidx,
                          hipStream_t st);

}  // namespace socu
