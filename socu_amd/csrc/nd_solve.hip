// ND forward and backward substitution: one kernel launch per level.
//
// Forward (levels ascending; reference K9-K12, block_tridiag_solver.py:425-578):
//   y_i <- L_i^{-1} y_i;  y_{i+s} -= R y_i;  y_{i-s} -= Lc^T y_i.
// The reference scatters the two updates with atomics.  Here a "forward
// step" for level l runs one workgroup per receiving block y_j (j = 2s(u+1)-1),
// which gathers both of level l's contributions,
//   y_j -= R_{j-s} y_{j-s} + Lc_{j+s}^T y_{j+s},
// and, when j is a node of level l+1, immediately applies y_j <- L_j^{-1} y_j.
// Step "-1" only solves the level-0 nodes.  No atomics, deterministic.
//
// Backward (levels descending; reference K13-K16, :581-736):
//   x_i -= R_i^T x_{i+s} + Lc_i x_{i-s};  x_i <- L_i^{-T} x_i.
// Each node owns x_i, so one workgroup per (node, column chunk).
#include "nd_common.h"
#include "nd_internal.h"

namespace socu {

// Packed tril(L_j) and reciprocal pivots into LDS.
template <typename T>
__device__ __forceinline__ void load_packed_tril(T* __restrict__ P, T* __restrict__ dinv, const T* __restrict__ A,
                                                 int n) {
  for (int idx = threadIdx.x; idx < n * n; idx += kThreads) {
    const int r = idx / n, c = idx - r * n;
    if (c <= r) P[tri(r) + c] = A[idx];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += kThreads) dinv[k] = T(1) / P[tri(k) + k];
  __syncthreads();
}

// W (n x w, row pitch WP) <- L^{-1} W, column sweep, one barrier per step.
template <typename T>
__device__ __forceinline__ void lower_solve_lds(T* __restrict__ W, const T* __restrict__ P,
                                                const T* __restrict__ dinv, int n, int w, int WP) {
  for (int k = 0; k < n - 1; ++k) {
    const T dk = dinv[k];
    const int rows = n - k - 1;
    for (int q = threadIdx.x; q < rows * w; q += kThreads) {
      const int rr = q / w, j = q - rr * w;
      const int r = k + 1 + rr;
      W[r * WP + j] -= P[tri(r) + k] * (W[k * WP + j] * dk);
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < n * w; q += kThreads) {
    const int r = q / w, j = q - r * w;
    W[r * WP + j] *= dinv[r];
  }
  __syncthreads();
}

// W <- L^{-T} W, backward column sweep: at step k rows r < k subtract L_kr z_k.
template <typename T>
__device__ __forceinline__ void upper_solve_T_lds(T* __restrict__ W, const T* __restrict__ P,
                                                  const T* __restrict__ dinv, int n, int w, int WP) {
  for (int k = n - 1; k > 0; --k) {
    const T dk = dinv[k];
    for (int q = threadIdx.x; q < k * w; q += kThreads) {
      const int r = q / w, j = q - r * w;
      W[r * WP + j] -= P[tri(k) + r] * (W[k * WP + j] * dk);
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < n * w; q += kThreads) {
    const int r = q / w, j = q - r * w;
    W[r * WP + j] *= dinv[r];
  }
  __syncthreads();
}

// acc(r, j) over the workgroup's outputs += sum_k A(r,k) Y[k][j]; A is an n x n
// row-major block read either as is (trans=false) or transposed (trans=true).
// A is staged through LDS in 32-wide K chunks so global reads stay contiguous.
template <typename T>
__device__ __forceinline__ void block_times_panel(T* __restrict__ W, const T* __restrict__ A, bool trans,
                                                  const T* __restrict__ Y, T* __restrict__ S, int n, int w, int WP) {
  // S: n x (kKChunk+1) staging of A(:, k0:k0+32)
  constexpr int SP = kKChunk + 1;
  for (int k0 = 0; k0 < n; k0 += kKChunk) {
    const int kc = min(kKChunk, n - k0);
    if (!trans) {
      for (int idx = threadIdx.x; idx < n * kKChunk; idx += kThreads) {
        const int r = idx / kKChunk, kk = idx - r * kKChunk;
        S[r * SP + kk] = (kk < kc) ? A[(size_t)r * n + k0 + kk] : T(0);
      }
    } else {
      // A^T(r, k) = A[k][r]: read rows k of A, contiguous in r
      for (int idx = threadIdx.x; idx < kKChunk * n; idx += kThreads) {
        const int kk = idx / n, r = idx - kk * n;
        if (kk < kc) S[r * SP + kk] = A[(size_t)(k0 + kk) * n + r];
        else S[r * SP + kk] = T(0);
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n * w; q += kThreads) {
      const int r = q / w, j = q - r * w;
      T acc = T(0);
      for (int kk = 0; kk < kc; ++kk) acc += S[r * SP + kk] * Y[(k0 + kk) * WP + j];
      W[r * WP + j] -= acc;
    }
    __syncthreads();
  }
}

// Load / store an n x w column chunk of a right-hand-side block.
template <typename T>
__device__ __forceinline__ void load_panel(T* __restrict__ W, const T* __restrict__ X, int n, int m, int c0, int w,
                                           int WP) {
  for (int q = threadIdx.x; q < n * w; q += kThreads) {
    const int r = q / w, j = q - r * w;
    W[r * WP + j] = X[(size_t)r * m + c0 + j];
  }
}
template <typename T>
__device__ __forceinline__ void store_panel(T* __restrict__ X, const T* __restrict__ W, int n, int m, int c0, int w,
                                            int WP) {
  for (int q = threadIdx.x; q < n * w; q += kThreads) {
    const int r = q / w, j = q - r * w;
    X[(size_t)r * m + c0 + j] = W[r * WP + j];
  }
}

// LDS layout shared by both solve kernels:
//   W  (n x WP)  the block being updated / solved
//   Y  (n x WP)  a neighbour block
//   S  (n x 33)  staging of one E block chunk; later reused for packed tril(L)
//   P, dinv      packed tril(L) (n(n+1)/2) and 1/L_kk (n), placed after S when
//                they do not fit in it.
__host__ __device__ inline size_t solve_lds_elems(int n) {
  constexpr int WP = kRhsChunk + 1;
  const size_t panel = (size_t)n * WP;
  const size_t stage = (size_t)n * (kKChunk + 1);
  const size_t tri_ = (size_t)n * (n + 1) / 2 + n;
  return 2 * panel + (stage > tri_ ? stage : tri_);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_forward_step(const T* __restrict__ L, const T* __restrict__ E,
                                                          T* __restrict__ x, const uint8_t* __restrict__ mask,
                                                          int N, int n, int m, int nE, int s, int off, int first) {
  // first != 0: step "-1", items are the level-0 nodes j = 2u, solve only.
  // otherwise:  items j = 2s(u+1)-1 receive level l's updates (stride s);
  //             j is a level-(l+1) node when u is even.
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int u = blockIdx.x;
  const int c0 = blockIdx.y * kRhsChunk;
  const int w = min(kRhsChunk, m - c0);
  constexpr int WP = kRhsChunk + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* W = reinterpret_cast<T*>(smem_raw);
  T* Y = W + (size_t)n * WP;
  T* S = Y + (size_t)n * WP;
  const size_t nn = (size_t)n * n;
  const T* Eb = E + (size_t)b * nE * nn;
  T* xb = x + (size_t)b * N * n * m;

  int j;
  bool do_solve;
  if (first) {
    j = 2 * u;
    do_solve = true;
  } else {
    j = 2 * s * (u + 1) - 1;
    do_solve = (u % 2) == 0;
  }
  load_panel(W, xb + (size_t)j * n * m, n, m, c0, w, WP);
  if (!first) {
    // left neighbour j-s (node of level l): y_j -= R y_{j-s}, R = block (j, j-s)
    __syncthreads();
    load_panel(Y, xb + (size_t)(j - s) * n * m, n, m, c0, w, WP);
    __syncthreads();
    block_times_panel(W, Eb + (size_t)(off + (j - s) / s) * nn, false, Y, S, n, w, WP);
    if (j + s < N) {
      // right neighbour j+s: y_j -= Lc^T y_{j+s}, Lc = block (j+s, j)
      load_panel(Y, xb + (size_t)(j + s) * n * m, n, m, c0, w, WP);
      __syncthreads();
      block_times_panel(W, Eb + (size_t)(off + j / s) * nn, true, Y, S, n, w, WP);
    }
  }
  if (do_solve) {
    T* P = S;
    T* dinv = P + tri(n);
    __syncthreads();
    load_packed_tril(P, dinv, L + ((size_t)b * N + j) * nn, n);
    lower_solve_lds(W, P, dinv, n, w, WP);
  } else {
    __syncthreads();
  }
  store_panel(xb + (size_t)j * n * m, W, n, m, c0, w, WP);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_backward_level(const T* __restrict__ L, const T* __restrict__ E,
                                                            T* __restrict__ x, const uint8_t* __restrict__ mask,
                                                            int N, int n, int m, int nE, int s, int off) {
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  const int c0 = blockIdx.y * kRhsChunk;
  const int w = min(kRhsChunk, m - c0);
  constexpr int WP = kRhsChunk + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* W = reinterpret_cast<T*>(smem_raw);
  T* Y = W + (size_t)n * WP;
  T* S = Y + (size_t)n * WP;
  const size_t nn = (size_t)n * n;
  const T* Eb = E + (size_t)b * nE * nn;
  T* xb = x + (size_t)b * N * n * m;

  load_panel(W, xb + (size_t)i * n * m, n, m, c0, w, WP);
  if (i + s < N) {
    // x_i -= R^T x_{i+s}, R = block (i+s, i)
    load_panel(Y, xb + (size_t)(i + s) * n * m, n, m, c0, w, WP);
    __syncthreads();
    block_times_panel(W, Eb + (size_t)(off + i / s) * nn, true, Y, S, n, w, WP);
  }
  if (i >= s) {
    // x_i -= Lc x_{i-s}, Lc = block (i, i-s)
    __syncthreads();
    load_panel(Y, xb + (size_t)(i - s) * n * m, n, m, c0, w, WP);
    __syncthreads();
    block_times_panel(W, Eb + (size_t)(off + (i - s) / s) * nn, false, Y, S, n, w, WP);
  }
  T* P = S;
  T* dinv = P + tri(n);
  __syncthreads();
  load_packed_tril(P, dinv, L + ((size_t)b * N + i) * nn, n);
  upper_solve_T_lds(W, P, dinv, n, w, WP);
  store_panel(xb + (size_t)i * n * m, W, n, m, c0, w, WP);
}

template <typename T>
static hipError_t set_solve_lds_limits() {
  const int bytes = (int)(solve_lds_elems(kMaxN) * sizeof(T));
  hipError_t e = hipFuncSetAttribute((const void*)k_forward_step<T>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_backward_level<T>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

hipError_t init_solve_kernels() {
  hipError_t e = set_solve_lds_limits<float>();
  if (e != hipSuccess) return e;
  return set_solve_lds_limits<double>();
}

// Forward step `step` in [-1, Lv-2]: step -1 solves level 0; step l gathers
// level l's updates and solves level l+1.
template <typename T>
hipError_t forward_step(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, int step,
                        hipStream_t st) {
  const int nE = (int)off_diag_storage_len(N);
  const int mch = (m + kRhsChunk - 1) / kRhsChunk;
  const size_t lds = solve_lds_elems(n) * sizeof(T);
  int items, s = 0, off = 0;
  if (step < 0) {
    items = (N + 1) / 2;
  } else {
    const Level lv = level_info(N, step);
    s = lv.s;
    off = lv.off;
    items = N / (2 * s);
  }
  if (items <= 0 || B <= 0 || mch <= 0) return hipSuccess;
  kt_begin(kKcForward, st);
  hipLaunchKernelGGL(k_forward_step<T>, dim3(items, mch, B), dim3(kThreads), lds, st, L, E, x, mask, N, n, m, nE, s,
                     off, step < 0 ? 1 : 0);
  kt_end(kKcForward, st);
  return hipGetLastError();
}

template <typename T>
hipError_t backward_level(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, int lidx,
                          hipStream_t st) {
  const int nE = (int)off_diag_storage_len(N);
  const int mch = (m + kRhsChunk - 1) / kRhsChunk;
  const size_t lds = solve_lds_elems(n) * sizeof(T);
  const Level lv = level_info(N, lidx);
  if (lv.dim <= 0 || B <= 0 || mch <= 0) return hipSuccess;
  kt_begin(kKcBackward, st);
  hipLaunchKernelGGL(k_backward_level<T>, dim3(lv.dim, mch, B), dim3(kThreads), lds, st, L, E, x, mask, N, n, m, nE,
                     lv.s, lv.off);
  kt_end(kKcBackward, st);
  return hipGetLastError();
}

template hipError_t forward_step<float>(const float*, const float*, float*, const uint8_t*, int, int, int, int, int,
                                        hipStream_t);
template hipError_t forward_step<double>(const double*, const double*, double*, const uint8_t*, int, int, int, int,
                                         int, hipStream_t);
template hipError_t backward_level<float>(const float*, const float*, float*, const uint8_t*, int, int, int, int, int,
                                          hipStream_t);
template hipError_t backward_level<double>(const double*, const double*, double*, const uint8_t*, int, int, int, int,
                                           int, hipStream_t);

}  // namespace socu
