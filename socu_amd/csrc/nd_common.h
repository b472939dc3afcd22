// Shared definitions for the nested-dissection (ND) block-tridiagonal
// Cholesky kernels on gfx950.
//
// Data layout in HBM (identical to socu's, README.md:54-80):
//   L  (B, N,   n, n) row-major, diagonal blocks; only tril is read.
//   E  (B, n_E, n, n) row-major, level segments of the ND fill-in storage;
//      segment l starts at off_l = sum_{l'<l} (N >> l') - 1 and entry k couples
//      column (k+1)s-1 with row (k+2)s-1, s = 2^l
//      (block_tridiag_solver.py:18-23, :793-813).
//   x  (B, N,   n, m) row-major right-hand sides / solutions.
//   mask (B,) bytes, optional; a zero entry leaves that system untouched.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace socu {

// Threads per workgroup for every level kernel: four 64-wide wavefronts.
constexpr int kThreads = 256;

// Largest block size n handled by the LDS-resident kernels (a padded
// 128x129 fp64 block is 129 KiB of the 160 KiB of LDS per workgroup).
constexpr int kMaxN = 128;

// Largest block size n of the tiled kernels for larger blocks (the block in
// HBM, 32 x 32 tiles through LDS; a staged 32-wide chunk of a 256-wide
// block is 66 KiB of FP64 LDS).
constexpr int kMaxNLarge = 256;

// Column chunk of the right-hand sides processed by one workgroup.
constexpr int kRhsChunk = 16;

struct Level {
  int l;         // level index
  int s;         // stride 2^l
  int dim;       // nodes at this level: i = s-1+2*s*t, t < dim
  int off;       // first E index of this level's segment
  int next_off;  // first E index of level l+1's segment
};

// Ties a loaded value to a register at this point of the program, so the
// compiler keeps the load unconditional and issued with its neighbours
// instead of sinking it into a later conditional use.
__device__ __forceinline__ void pin_reg(float& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void pin_reg(double& v) { asm volatile("" : "+v"(v)); }

__host__ __device__ inline int recursive_iterations(int N) {
  int L = 0;
  while (N > 0) { ++L; N >>= 1; }
  return L;
}

__host__ __device__ inline long long off_diag_storage_len(int N) {
  long long t = 0;
  const int L = recursive_iterations(N);
  for (int l = 0; l < L; ++l) t += (N >> l) - 1;
  return t;
}

__host__ inline Level level_info(int N, int l) {
  Level v;
  v.l = l;
  v.s = 1 << l;
  v.dim = (N + v.s) / (2 * v.s);
  int off = 0;
  for (int q = 0; q < l; ++q) off += (N >> q) - 1;
  v.off = off;
  v.next_off = off + (N >> l) - 1;
  return v;
}

// v_rsq_f32 directly: rsqrtf adds a rescaling for denormal arguments around
// the same instruction (identical results for normal arguments); a pivot
// below 2^-126 means a numerically singular block anyway.
__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }
// v_rsq_f64 (an estimate) refined by one third-order step (below), which
// triples the correct bits.  The pivots here are positive and normal;
// rsqrt(double) spends six selects per call on zero, infinite and denormal
// arguments, on the pivot chain.  A non-positive or NaN pivot still gives
// NaN (zero: inf * 0).
__device__ __forceinline__ double rsqrt_t(double x) {
  // one third-order correction of the hardware estimate y: with
  // e = 1 - x y^2, 1/sqrt(x) = y (1 - e)^(-1/2) = y (1 + e/2 + 3e^2/8 + ...),
  // and the dropped terms are ~e^3 (e is ~2^-22 here)
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y, e * fma(e, 0.375, 0.5), y);
}
__device__ __forceinline__ float sqrt_t(float x) { return sqrtf(x); }
__device__ __forceinline__ double sqrt_t(double x) { return sqrt(x); }

// Agent-scope relaxed atomic loads and stores (global_load/store ... sc1): the
// L2 of an XCD writes such stores through and does not serve such loads
// stale, so values handed between workgroups of one kernel need no L2
// write-back or invalidation; stores_done() waits until this thread's stores
// have completed (they then precede its later stores and counter updates).
template <typename T>
__device__ __forceinline__ T ld_coh(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_coh(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stores_done() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0);
  asm volatile("" ::: "memory");
}

// True when batch entry b is switched off by the mask.
__device__ __forceinline__ bool masked_out(const uint8_t* mask, int b) {
  return mask != nullptr && mask[b] == 0;
}

}  // namespace socu
