// ND block-tridiagonal Cholesky factor: three kernels per dissection level.
//
//   k_potrf  one workgroup per node i:      L_i <- chol(L_i)
//            (replaces potrf_blocked, socu/blas/potrf.py:97-119, kernel K2
//             block_tridiag_solver.py:197-216)
//   k_trsm   one workgroup per (node, 32-wide chunk):
//            R  <- R L_i^{-T}   (trsm_rltn_blocked, trsm.py:94-120, K3 :219-248)
//            Lc <- L_i^{-1} Lc  (trsm_llnn_blocked, trsm.py:202-228, K4 :251-279)
//   k_schur  one workgroup per 32x32 output tile, owner-computes:
//            L_j -= R_{j-s} R_{j-s}^T + Lc_{j+s}^T Lc_{j+s}
//                   (syrk_ln / syrk_lt, syrk.py:106-158, :264-315, K5/K6)
//            F_i -= R_i Lc_i   (gemm_nn_blocked, gemm.py:151-201, K7 :345-377)
//
// The reference scatters the two syrk updates into L_{i+-s} with atomics
// (K5/K6).  Here each lower tile of a receiving block L_j is computed by one
// workgroup that sums both contributions, so there are no atomics and the
// result is bitwise deterministic run to run.
#include "nd_common.h"
#include "nd_internal.h"

namespace socu {

// ---------------------------------------------------------------------------
// potrf: right-looking Cholesky of one n x n block held as a packed lower
// triangle in LDS.  Step k subtracts a_rk a_ck / a_kk from every trailing
// element; the pivots are square-rooted and the columns scaled at the end.
// A non-positive pivot yields NaN silently, as in the reference.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_potrf(T* __restrict__ L, const uint8_t* __restrict__ mask,
                                                   int N, int n, int s) {
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  T* A = L + ((size_t)b * N + i) * (size_t)n * n;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* P = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x;

  for (int idx = tid; idx < n * n; idx += kThreads) {
    const int r = idx / n, c = idx - r * n;
    if (c <= r) P[tri(r) + c] = A[idx];
  }
  __syncthreads();

  for (int k = 0; k < n - 1; ++k) {
    const T inv = T(1) / P[tri(k) + k];
    const int m = n - k - 1;
    for (int q = tid; q < m * m; q += kThreads) {
      const int rr = q / m, cc = q - rr * m;
      if (cc <= rr) {
        const int r = k + 1 + rr, c = k + 1 + cc;
        P[tri(r) + c] -= P[tri(r) + k] * P[tri(c) + k] * inv;
      }
    }
    __syncthreads();
  }

  for (int idx = tid; idx < n * n; idx += kThreads) {
    const int r = idx / n, c = idx - r * n;
    if (c <= r) {
      const T d = P[tri(c) + c];
      A[idx] = (c == r) ? sqrt_t(d) : P[tri(r) + c] * rsqrt_t(d);
    }
  }
}

// ---------------------------------------------------------------------------
// trsm: solve L_i Z = W for an n x w panel W held in LDS (w <= kTrsmChunk).
// For the R task W is a chunk of rows of R, transposed; for the Lc task W is a
// chunk of columns of Lc.  Column sweep: at step k every row r > k subtracts
// L_rk z_k, where z_k = W_k / L_kk is final after step k-1.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_trsm(const T* __restrict__ L, T* __restrict__ E,
                                                  const uint8_t* __restrict__ mask, int N, int n, int nE,
                                                  int s, int off, int nchunk) {
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  const bool isR = (int)blockIdx.y < nchunk;
  const int ch = isR ? blockIdx.y : blockIdx.y - nchunk;
  if (isR && !(i + s < N)) return;
  if (!isR && !(i >= s)) return;

  const T* A = L + ((size_t)b * N + i) * (size_t)n * n;
  T* X = E + ((size_t)b * nE + off + (isR ? i / s : (i - s) / s)) * (size_t)n * n;
  const int c0 = ch * kTrsmChunk;
  const int w = min(kTrsmChunk, n - c0);
  constexpr int WP = kTrsmChunk + 1;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* P = reinterpret_cast<T*>(smem_raw);  // packed tril(L_i)
  T* dinv = P + tri(n);                    // 1 / L_kk
  T* W = dinv + n;                         // n x WP
  const int tid = threadIdx.x;

  for (int idx = tid; idx < n * n; idx += kThreads) {
    const int r = idx / n, c = idx - r * n;
    if (c <= r) P[tri(r) + c] = A[idx];
  }
  if (isR) {
    // W[k][j] = R[c0+j][k]
    for (int idx = tid; idx < w * n; idx += kThreads) {
      const int j = idx / n, k = idx - j * n;
      W[k * WP + j] = X[(size_t)(c0 + j) * n + k];
    }
  } else {
    // W[k][j] = Lc[k][c0+j]
    for (int idx = tid; idx < n * w; idx += kThreads) {
      const int k = idx / w, j = idx - k * w;
      W[k * WP + j] = X[(size_t)k * n + c0 + j];
    }
  }
  __syncthreads();
  for (int k = tid; k < n; k += kThreads) dinv[k] = T(1) / P[tri(k) + k];
  __syncthreads();

  for (int k = 0; k < n - 1; ++k) {
    const T dk = dinv[k];
    const int rows = n - k - 1;
    for (int q = tid; q < rows * w; q += kThreads) {
      const int rr = q / w, j = q - rr * w;
      const int r = k + 1 + rr;
      W[r * WP + j] -= P[tri(r) + k] * (W[k * WP + j] * dk);
    }
    __syncthreads();
  }

  if (isR) {
    for (int idx = tid; idx < w * n; idx += kThreads) {
      const int j = idx / n, k = idx - j * n;
      X[(size_t)(c0 + j) * n + k] = W[k * WP + j] * dinv[k];
    }
  } else {
    for (int idx = tid; idx < n * w; idx += kThreads) {
      const int k = idx / w, j = idx - k * w;
      X[(size_t)k * n + c0 + j] = W[k * WP + j] * dinv[k];
    }
  }
}

// ---------------------------------------------------------------------------
// Schur-complement updates of one level, owner-computes.
//   task < nT*ntri : lower tile (tr, tc) of receiving block L_j,
//                    j = 2s*(u+1)-1, u = task / ntri
//   otherwise      : tile (tr, tc) of fill-in F of node i = s-1+2s*t, t >= 1
// Each 32x32 tile is one workgroup: 256 threads x 2x2 outputs; K staged through
// LDS in chunks of 32.  Operand forms: R R^T (A B^T), Lc^T Lc (A^T B), R Lc (A B).
// ---------------------------------------------------------------------------
enum { kFormABt = 0, kFormAtB = 1, kFormAB = 2 };

template <typename T>
__device__ __forceinline__ void stage_operands(T (*As)[kTile + 1], T (*Bs)[kTile + 1], const T* __restrict__ Am,
                                               const T* __restrict__ Bm, int form, int n, int r0, int c0, int k0) {
  // As[kk][r] = A(r0+r, k0+kk), Bs[kk][c] = B(k0+kk, c0+c), zero outside n.
  // Two index mappings over the same 32x32 chunk: (a, kk) with kk fastest reads
  // a row-major operand along its rows; (kk, a) with a fastest reads it along
  // its columns.  Each operand uses the mapping that keeps reads contiguous.
  static_assert(kTile == kKChunk, "square staging chunk");
  const int tid = threadIdx.x;
  for (int idx = tid; idx < kTile * kKChunk; idx += kThreads) {
    const int a = idx / kKChunk, kk = idx - a * kKChunk, k = k0 + kk;
    const int kk2 = idx / kTile, a2 = idx - kk2 * kTile, k2 = k0 + kk2;
    if (form == kFormABt) {  // A(r,k) = Am[r][k], B(k,c) = Bm[c][k]
      As[kk][a] = (r0 + a < n && k < n) ? Am[(size_t)(r0 + a) * n + k] : T(0);
      Bs[kk][a] = (c0 + a < n && k < n) ? Bm[(size_t)(c0 + a) * n + k] : T(0);
    } else if (form == kFormAtB) {  // A(r,k) = Am[k][r], B(k,c) = Bm[k][c]
      As[kk2][a2] = (r0 + a2 < n && k2 < n) ? Am[(size_t)k2 * n + r0 + a2] : T(0);
      Bs[kk2][a2] = (c0 + a2 < n && k2 < n) ? Bm[(size_t)k2 * n + c0 + a2] : T(0);
    } else {  // A(r,k) = Am[r][k], B(k,c) = Bm[k][c]
      As[kk][a] = (r0 + a < n && k < n) ? Am[(size_t)(r0 + a) * n + k] : T(0);
      Bs[kk2][a2] = (c0 + a2 < n && k2 < n) ? Bm[(size_t)k2 * n + c0 + a2] : T(0);
    }
  }
}

template <typename T>
__device__ __forceinline__ void tile_product(T acc[2][2], const T* __restrict__ Am, const T* __restrict__ Bm, int form,
                                             int n, int r0, int c0, T (*As)[kTile + 1], T (*Bs)[kTile + 1]) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  for (int k0 = 0; k0 < n; k0 += kKChunk) {
    __syncthreads();
    stage_operands<T>(As, Bs, Am, Bm, form, n, r0, c0, k0);
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < kKChunk; ++kk) {
      const T a0 = As[kk][2 * ty], a1 = As[kk][2 * ty + 1];
      const T b0 = Bs[kk][2 * tx], b1 = Bs[kk][2 * tx + 1];
      acc[0][0] += a0 * b0;
      acc[0][1] += a0 * b1;
      acc[1][0] += a1 * b0;
      acc[1][1] += a1 * b1;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_schur(T* __restrict__ L, T* __restrict__ E,
                                                   const uint8_t* __restrict__ mask, int N, int n, int nE, int s,
                                                   int off, int next_off, int nT, int ntile) {
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  __shared__ T As[kKChunk][kTile + 1];
  __shared__ T Bs[kKChunk][kTile + 1];
  const int ntri = ntile * (ntile + 1) / 2;
  int task = blockIdx.x;
  T acc[2][2] = {{T(0), T(0)}, {T(0), T(0)}};
  T* Eb = E + (size_t)b * nE * n * n;
  const size_t nn = (size_t)n * n;
  T* C;
  int tr, tc;
  if (task < nT * ntri) {
    const int u = task / ntri;
    int q = task - u * ntri;
    tr = 0;
    while (q > tr) { q -= tr + 1; ++tr; }
    tc = q;
    const int j = 2 * s * (u + 1) - 1;
    C = L + ((size_t)b * N + j) * nn;
    const T* R = Eb + (size_t)(off + (j - s) / s) * nn;  // block (j, j-s)
    tile_product<T>(acc, R, R, kFormABt, n, tr * kTile, tc * kTile, As, Bs);
    if (j + s < N) {
      const T* Lc = Eb + (size_t)(off + j / s) * nn;  // block (j+s, j)
      tile_product<T>(acc, Lc, Lc, kFormAtB, n, tr * kTile, tc * kTile, As, Bs);
    }
  } else {
    task -= nT * ntri;
    const int f = task / (ntile * ntile);
    const int q = task - f * ntile * ntile;
    tr = q / ntile;
    tc = q - tr * ntile;
    const int t = f + 1;
    const int i = s - 1 + 2 * s * t;
    const T* R = Eb + (size_t)(off + i / s) * nn;
    const T* Lc = Eb + (size_t)(off + (i - s) / s) * nn;
    C = Eb + (size_t)(next_off + (i - s) / (2 * s)) * nn;
    tile_product<T>(acc, R, Lc, kFormAB, n, tr * kTile, tc * kTile, As, Bs);
  }
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int r = tr * kTile + 2 * ty + a, cc = tc * kTile + 2 * tx + c;
      if (r < n && cc < n) C[(size_t)r * n + cc] -= acc[a][c];
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static size_t potrf_lds(int n, size_t w) { return w * (size_t)(n * (n + 1) / 2); }
static size_t trsm_lds(int n, size_t w) {
  return w * ((size_t)(n * (n + 1) / 2) + n + (size_t)n * (kTrsmChunk + 1));
}

template <typename T>
static hipError_t set_factor_lds_limits() {
  hipError_t e = hipFuncSetAttribute((const void*)k_potrf<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)potrf_lds(kMaxN, sizeof(T)));
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_trsm<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)trsm_lds(kMaxN, sizeof(T)));
}

hipError_t init_factor_kernels() {
  hipError_t e = set_factor_lds_limits<float>();
  if (e != hipSuccess) return e;
  return set_factor_lds_limits<double>();
}

template <typename T>
hipError_t factor_level(T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, hipStream_t st) {
  const int nE = (int)off_diag_storage_len(N);
  const int Lv = recursive_iterations(N);
  const Level lv = level_info(N, lidx);
  if (lv.dim <= 0 || B <= 0) return hipSuccess;

  kt_begin(kKcPotrf, st);
  hipLaunchKernelGGL(k_potrf<T>, dim3(lv.dim, B), dim3(kThreads), potrf_lds(n, sizeof(T)), st, L, mask, N, n, lv.s);
  kt_end(kKcPotrf, st);
  if (lidx == Lv - 1) return hipGetLastError();

  const int nchunk = (n + kTrsmChunk - 1) / kTrsmChunk;
  kt_begin(kKcTrsm, st);
  hipLaunchKernelGGL(k_trsm<T>, dim3(lv.dim, 2 * nchunk, B), dim3(kThreads), trsm_lds(n, sizeof(T)), st, L, E, mask, N,
                     n, nE, lv.s, lv.off, nchunk);
  kt_end(kKcTrsm, st);

  const int ntile = (n + kTile - 1) / kTile;
  const int ntri = ntile * (ntile + 1) / 2;
  const int nT = N / (2 * lv.s);
  int nF = 0;
  for (int t = 1; t < lv.dim; ++t)
    if (lv.s - 1 + 2 * lv.s * t + lv.s < N) ++nF;
  // fill-in nodes are t = 1..nF (contiguous: i+s<N fails only at the tail)
  const int tasks = nT * ntri + nF * ntile * ntile;
  if (tasks > 0) {
    kt_begin(kKcSchur, st);
    hipLaunchKernelGGL(k_schur<T>, dim3(tasks, B), dim3(kThreads), 0, st, L, E, mask, N, n, nE, lv.s, lv.off,
                       lv.next_off, nT, ntile);
    kt_end(kKcSchur, st);
  }
  return hipGetLastError();
}

template hipError_t factor_level<float>(float*, float*, const uint8_t*, int, int, int, int, hipStream_t);
template hipError_t factor_level<double>(double*, double*, const uint8_t*, int, int, int, int, hipStream_t);

}  // namespace socu
