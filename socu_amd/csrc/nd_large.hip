// Block sizes kMaxN < n <= kMaxNLarge: the per-node kernels of the ND factor
// and sweeps for blocks too large to stay in LDS (a 256 x 257 FP64 block is
// 514 KiB; a workgroup has 160 KiB).  The block stays in HBM / L2 and passes
// through LDS one 32 x 32 tile at a time; every product is a 32 x 32 MFMA
// tile product of one wavefront.  Same arithmetic as the LDS-resident
// kernels of nd_blocked.hip, in the same storage:
//
//   k_potrf_large  right-looking blocked Cholesky of L_i in place (the
//                  reference's potrf_blocked, socu/blas/potrf.py:97-119), then
//                  the transposed inverse Linv^T written into the strict
//                  upper triangle (the layout k_potrf_blk leaves), band by
//                  band: Linv_ij = -Dinv_i sum_{k=j}^{i-1} L_ik Linv_kj.
//   k_trtri_large  the same inverse from tril(L) alone (a factor made
//                  anywhere), into a workspace W.
//   k_trimul_large products with U = Linv^T as stored: R <- R U (trsm_rltn,
//                  trsm.py:94-120), Lc <- U^T Lc (trsm_llnn, trsm.py:202-228),
//                  y <- U^T y (the forward apply) and x <- U x (the backward
//                  apply, trsm_lltn, trsm.py:314-340).  A workgroup owns a
//                  32-row (R) or 32-column chunk of its operand and stages it
//                  in LDS before writing any of it, so the in-place update has
//                  no race between workgroups.
//   k_update_large the sweep updates y_j -= R y_{j-s} + Lc^T y_{j+s} and
//                  x_i -= R_i^T x_{i+s} + Lc_i x_{i-s} (gemm_nn / gemm_tn,
//                  gemm.py:151-201, :358-408), one wavefront per 32 x 32
//                  output tile.
// The Schur updates use k_schur_blk, which already walks any n in chunks.
#include "../../include/socu_hip.h"
#include "nd_common.h"
#include "nd_internal.h"
#include "nd_tile.h"

namespace socu {

namespace {

constexpr int kLP = 33;                       // pitch of an LDS tile
constexpr int kTileElems = 32 * kLP;
constexpr int kLargeTiles = kMaxNLarge / 32;  // tiles per block dimension, at most

// LDS operations of one wavefront complete in issue order; this keeps the
// compiler from moving LDS accesses of the wave across the point.
__device__ __forceinline__ void wave_lds_order() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Tile (r0, c0) of a rows x cols row-major matrix M (leading dimension ld)
// into the wave's LDS tile Tl: entries outside the matrix read 0, or 1 on
// the matrix diagonal when pad_diag (identity padding of a diagonal block).
// Lane l moves row 2q + l/32, column l%32; loads are unconditional from
// clamped addresses.
template <typename T>
__device__ __forceinline__ void wtile_load(T* Tl, const T* __restrict__ M, int rows, int cols, int ld, int r0, int c0,
                                           bool pad_diag) {
  const int l = lane_id(), lr = l >> 5, lc = l & 31;
  T v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = r0 + 2 * q + lr, c = c0 + lc;
    v[q] = M[(size_t)min(r, rows - 1) * ld + min(c, cols - 1)];
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = r0 + 2 * q + lr, c = c0 + lc;
    const bool in = r < rows && c < cols;
    Tl[(2 * q + lr) * kLP + lc] = in ? v[q] : ((pad_diag && r == c) ? T(1) : T(0));
  }
  wave_lds_order();
}

// Tl <- the transpose of tile (r0, c0) of M: Tl[a][b] = M[r0 + b][c0 + a].
template <typename T>
__device__ __forceinline__ void wtile_load_t(T* Tl, const T* __restrict__ M, int rows, int cols, int ld, int r0,
                                             int c0) {
  const int l = lane_id(), lr = l >> 5, lc = l & 31;
  T v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = r0 + 2 * q + lr, c = c0 + lc;
    v[q] = M[(size_t)min(r, rows - 1) * ld + min(c, cols - 1)];
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = r0 + 2 * q + lr, c = c0 + lc;
    Tl[lc * kLP + 2 * q + lr] = (r < rows && c < cols) ? v[q] : T(0);
  }
  wave_lds_order();
}

// Tile (kb, q) of U = Linv^T of a diagonal block A (n x n) as the factor
// leaves it: kb < q: A's tile (kb, q); kb == q: the strict upper triangle of
// A's diagonal tile, 1 / A[r][r] on the diagonal, zeros below; identity
// outside n.
template <typename T>
__device__ __forceinline__ void wtile_load_u(T* Tl, const T* __restrict__ A, int n, int kb, int q) {
  if (kb != q) {
    wtile_load(Tl, A, n, n, n, 32 * kb, 32 * q, false);
    return;
  }
  const int l = lane_id(), lr = l >> 5, lc = l & 31;
  T v[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int r = 32 * q + 2 * e + lr, c = 32 * q + lc;
    v[e] = A[(size_t)min(r, n - 1) * n + min(c, n - 1)];
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int rr = 2 * e + lr, r = 32 * q + rr, c = 32 * q + lc;
    T u;
    if (r < n && c < n) u = (rr < lc) ? v[e] : ((rr == lc) ? T(1) / v[e] : T(0));
    else u = (rr == lc) ? T(1) : T(0);
    Tl[rr * kLP + lc] = u;
  }
  wave_lds_order();
}

// Tl <- Linv_kj (k >= j) from a block whose strict upper triangle holds
// Linv^T: k > j: the transpose of tile (j, k); k == j: the strict upper
// triangle of diagonal tile j transposed, dinv (1 / L[r][r], 1 on padded
// rows) on the diagonal, zeros above.
template <typename T>
__device__ __forceinline__ void wtile_load_linv(T* Tl, const T* __restrict__ A, int n, int k, int j, const T* dinv) {
  if (k != j) {
    wtile_load_t(Tl, A, n, n, n, 32 * j, 32 * k);
    return;
  }
  const int l = lane_id(), lr = l >> 5, lc = l & 31;
  T v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = 32 * j + 2 * q + lr, c = 32 * j + lc;
    v[q] = A[(size_t)min(r, n - 1) * n + min(c, n - 1)];
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    // stored (row a = 2q + lr, column b = lc) holds Linv[b][a] for b > a
    const int a = 2 * q + lr, b = lc;
    const bool in = 32 * j + a < n && 32 * j + b < n;
    T u = T(0);
    if (b > a) u = in ? v[q] : T(0);
    Tl[b * kLP + a] = u;
  }
  wave_lds_order();
  if (l < 32) Tl[l * kLP + l] = dinv[32 * j + l];
  wave_lds_order();
}

// acc (32 x 32) into tile (r0, c0) of M: M = v (mode 0), M -= v (mode 1),
// entries inside rows x cols only.
template <typename T>
__device__ __forceinline__ void acc_to_global(typename Mfma<T>::acc_t acc[2][2], T* __restrict__ M, int rows, int cols,
                                              int ld, int r0, int c0, int mode) {
  acc_foreach<T>(acc, [&](int r, int c, T v) {
    const int rr = r0 + r, cc = c0 + c;
    if (rr < rows && cc < cols) {
      T* p = M + (size_t)rr * ld + cc;
      *p = mode == 0 ? v : *p - v;
    }
  });
}

template <typename T>
__device__ __forceinline__ void acc_to_tile(typename Mfma<T>::acc_t acc[2][2], T* Tl) {
  acc_foreach<T>(acc, [&](int r, int c, T v) { Tl[r * kLP + c] = v; });
  wave_lds_order();
}

template <typename T>
__device__ __forceinline__ void acc_clear(typename Mfma<T>::acc_t acc[2][2]) {
  acc[0][0] = acc[0][1] = acc[1][0] = acc[1][1] = acc_zero<T>();
}

// In-bounds entries of LDS tile Tl into tile (t0, t0) of A (n x n).
template <typename T>
__device__ __forceinline__ void wtile_store_diag(const T* Tl, T* __restrict__ A, int n, int t0) {
  const int l = lane_id(), lr = l >> 5, lc = l & 31;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = t0 + 2 * q + lr, c = t0 + lc;
    if (r < n && c < n) A[(size_t)r * n + c] = Tl[(2 * q + lr) * kLP + lc];
  }
}

// Inverse bands of a factored block A (lower: L; diagonal tiles' strict
// upper triangles: their Linv^T; dinv in LDS): the off-diagonal tiles of
// Linv^T into A's upper tiles, band d = i - j after band d - 1.
template <typename T>
__device__ __forceinline__ void inverse_bands(T* __restrict__ A, int n, int NB, const T* dinv, T* X0, T* X1) {
  const int wave = wave_id();
  for (int d = 1; d < NB; ++d) {
    for (int j = wave; j + d < NB; j += 4) {
      const int i = j + d;
      typename Mfma<T>::acc_t acc[2][2];
      acc_clear<T>(acc);
      for (int k = j; k < i; ++k) {  // S = sum_k L_ik Linv_kj
        wtile_load(X0, A, n, n, n, 32 * i, 32 * k, false);
        wtile_load_linv(X1, A, n, k, j, dinv);
        wave_gemm<T>(acc, X0, kLP, 1, X1, kLP, 1, 32);
      }
      acc_to_tile<T>(acc, X1);
      wtile_load_linv(X0, A, n, i, i, dinv);  // Dinv_i = Linv_ii
      typename Mfma<T>::acc_t inv[2][2];
      acc_clear<T>(inv);
      wave_gemm<T>(inv, X0, kLP, 1, X1, kLP, 1, 32);
      // Linv_ij = -Dinv_i S, stored transposed in tile (j, i)
      acc_foreach<T>(inv, [&](int r, int c, T v) {
        const int rr = 32 * j + c, cc = 32 * i + r;
        if (rr < n && cc < n) A[(size_t)rr * n + cc] = -v;
      });
    }
    __syncthreads();
  }
}

template <typename T>
constexpr size_t potrf_large_lds() {
  // diagonal tile, ut, 128 column entries, dinv, two tiles per wavefront
  return sizeof(T) * (2 * kTileElems + 128 + kMaxNLarge + 8 * kTileElems);
}

// One workgroup of four wavefronts per node of the level (grid (dim, B)).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_potrf_large(T* __restrict__ L, const uint8_t* __restrict__ mask, int N,
                                                         int n, int s) {
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  T* A = L + ((size_t)b * N + i) * (size_t)n * n;
  const int NB = (n + 31) / 32;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* D = reinterpret_cast<T*>(smem_raw);
  T* ut = D + kTileElems;
  T* col = ut + kTileElems;
  T* dinv = col + 128;
  const int wave = wave_id();
  T* X0 = dinv + kMaxNLarge + 2 * wave * kTileElems;
  T* X1 = X0 + kTileElems;
  for (int kb = 0; kb < NB; ++kb) {
    if (wave == 0) {
      wtile_load(D, A, n, n, n, 32 * kb, 32 * kb, true);
      tile_chol_inv<T>(D, kLP, dinv + 32 * kb, col, ut);
      wave_lds_order();
      wtile_store_diag(D, A, n, 32 * kb);  // L below, Linv^T above the diagonal
    }
    __syncthreads();
    // panel: tile (ti, kb) <- tile (ti, kb) Linv_kb^T, with ut[k][c] = Linv[c][k]
    for (int ti = kb + 1 + wave; ti < NB; ti += 4) {
      wtile_load(X0, A, n, n, n, 32 * ti, 32 * kb, false);
      typename Mfma<T>::acc_t acc[2][2];
      acc_clear<T>(acc);
      wave_gemm<T>(acc, X0, kLP, 1, ut, kLP, 1, 32);
      acc_to_global<T>(acc, A, n, n, n, 32 * ti, 32 * kb, 0);
    }
    __syncthreads();
    // trailing update of the lower tiles (gi, gj), kb < gj <= gi
    const int mt = NB - kb - 1, cnt = mt * (mt + 1) / 2;
    for (int t = wave; t < cnt; t += 4) {
      int gi = 0, q = t;
      while (q > gi) {
        q -= gi + 1;
        ++gi;
      }
      gi += kb + 1;
      const int gj = q + kb + 1;
      wtile_load(X0, A, n, n, n, 32 * gi, 32 * kb, false);
      wtile_load(X1, A, n, n, n, 32 * gj, 32 * kb, false);
      typename Mfma<T>::acc_t acc[2][2];
      acc_clear<T>(acc);
      wave_gemm<T>(acc, X0, kLP, 1, X1, 1, kLP, 32);  // X(gi) X(gj)^T
      acc_to_global<T>(acc, A, n, n, n, 32 * gi, 32 * gj, 1);
    }
    __syncthreads();
  }
  inverse_bands<T>(A, n, NB, dinv, X0, X1);
}

template <typename T>
constexpr size_t trtri_large_lds() {
  return sizeof(T) * (kMaxNLarge + 8 * kTileElems);
}

// W_i <- tril(L_i) with the transposed inverse in its strict upper triangle,
// from tril(L_i) alone (grid (N, B)).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_trtri_large(const T* __restrict__ L, T* __restrict__ W,
                                                         const uint8_t* __restrict__ mask, int N, int n) {
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  const size_t nn = (size_t)n * n;
  const T* A = L + ((size_t)b * N + blockIdx.x) * nn;
  T* Wb = W + ((size_t)b * N + blockIdx.x) * nn;
  const int NB = (n + 31) / 32;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* dinv = reinterpret_cast<T*>(smem_raw);
  const int wave = wave_id();
  T* X0 = dinv + kMaxNLarge + 2 * wave * kTileElems;
  T* X1 = X0 + kTileElems;
  // strictly lower tiles: copied
  for (size_t idx = threadIdx.x; idx < nn; idx += kThreads) {
    const int r = (int)(idx / n), c = (int)(idx - (size_t)r * n);
    if (r / 32 > c / 32) Wb[idx] = A[idx];
  }
  // diagonal tiles: L below, its inverse transposed above, dinv
  for (int jb = wave; jb < NB; jb += 4) {
    wtile_load(X0, A, n, n, n, 32 * jb, 32 * jb, true);
    tile_tri_inv<T>(X0, kLP, dinv + 32 * jb);
    wave_lds_order();
    wtile_store_diag(X0, Wb, n, 32 * jb);
  }
  __syncthreads();
  inverse_bands<T>(Wb, n, NB, dinv, X0, X1);
}

template <typename T>
constexpr size_t trimul_large_lds() {
  // the staged chunk (32 x (NP + 1) or NP x 33) and one tile per wavefront
  return sizeof(T) * ((size_t)kMaxNLarge * kLP + 4 * kTileElems);
}

enum { kTrsm = 0, kApplyFwd = 1, kApplyBwd = 2 };

// kind kTrsm (grid (dim, 2 NB, B)): chunk y < NB: R_i rows 32y..  <- R U;
//   y >= NB: Lc_i columns 32(y - NB)..  <- U^T Lc.
// kind kApplyFwd / kApplyBwd (grid (dim, ceil(m / 32), B)): x_i columns
//   32y..  <- U^T x_i / U x_i.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_trimul_large(const T* __restrict__ L, T* __restrict__ E,
                                                          T* __restrict__ x, const uint8_t* __restrict__ mask, int N,
                                                          int n, int m, int nE, int s, int off, int kind) {
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  const int NB = (n + 31) / 32;
  const size_t nn = (size_t)n * n;
  const T* U = L + ((size_t)b * N + i) * nn;
  T* X;
  int ld, cols, r0 = 0, c0 = 0;
  bool rowsR = false, upper = false;  // R U (row chunk) / U X (backward apply)
  if (kind == kTrsm) {
    const int y = blockIdx.y;
    if (y < NB) {
      if (i + s >= N) return;
      X = E + ((size_t)b * nE + off + i / s) * nn;
      rowsR = true;
      r0 = 32 * y;
    } else {
      if (i < s) return;
      X = E + ((size_t)b * nE + off + (i - s) / s) * nn;
      c0 = 32 * (y - NB);
    }
    ld = n;
    cols = n;
  } else {
    X = x + ((size_t)b * N + i) * (size_t)n * m;
    ld = m;
    cols = m;
    c0 = 32 * blockIdx.y;
    upper = kind == kApplyBwd;
  }
  if (c0 >= cols) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Xs = reinterpret_cast<T*>(smem_raw);
  const int wave = wave_id();
  T* Ut = Xs + (size_t)kMaxNLarge * kLP + wave * kTileElems;
  const int NP = 32 * NB;
  const int XP = NP + 1;
  // stage the chunk: R rows as Xs[r][k] (pitch NP + 1), columns as Xs[k][c]
  // (pitch 33); zero outside
  if (rowsR) {
    for (int idx = threadIdx.x; idx < 32 * NP; idx += kThreads) {
      const int r = idx / NP, k = idx - r * NP;
      const T v = X[(size_t)min(r0 + r, n - 1) * ld + min(k, n - 1)];
      Xs[r * XP + k] = (r0 + r < n && k < n) ? v : T(0);
    }
  } else {
    for (int idx = threadIdx.x; idx < NP * 32; idx += kThreads) {
      const int k = idx >> 5, c = idx & 31;
      const T v = X[(size_t)min(k, n - 1) * ld + min(c0 + c, cols - 1)];
      Xs[k * kLP + c] = (k < n && c0 + c < cols) ? v : T(0);
    }
  }
  __syncthreads();
  for (int q = wave; q < NB; q += 4) {
    typename Mfma<T>::acc_t acc[2][2];
    acc_clear<T>(acc);
    if (rowsR) {  // out[:, q] = sum_{kb <= q} X[:, kb] U(kb, q)
      for (int kb = 0; kb <= q; ++kb) {
        wtile_load_u(Ut, U, n, kb, q);
        wave_gemm<T>(acc, Xs + 32 * kb, XP, 1, Ut, kLP, 1, 32);
      }
      acc_to_global<T>(acc, X, n, cols, ld, r0, 32 * q, 0);
    } else if (!upper) {  // out[q, :] = sum_{kb <= q} U(kb, q)^T X[kb, :]
      for (int kb = 0; kb <= q; ++kb) {
        wtile_load_u(Ut, U, n, kb, q);
        wave_gemm<T>(acc, Ut, 1, kLP, Xs + 32 * kb * kLP, kLP, 1, 32);
      }
      acc_to_global<T>(acc, X, n, cols, ld, 32 * q, c0, 0);
    } else {  // out[q, :] = sum_{kb >= q} U(q, kb) X[kb, :]
      for (int kb = q; kb < NB; ++kb) {
        wtile_load_u(Ut, U, n, q, kb);
        wave_gemm<T>(acc, Ut, kLP, 1, Xs + 32 * kb * kLP, kLP, 1, 32);
      }
      acc_to_global<T>(acc, X, n, cols, ld, 32 * q, c0, 0);
    }
  }
}

// Sweep updates, one wavefront per 32 x 32 output tile (row tile rb of the
// target block, right-hand-side columns c0..), grid (items, ceil(NB / 4) *
// mch, B) with mch = ceil(m / 32).
//   forward (target j = 2s(u+1)-1): y_j -= R_{j-s} y_{j-s} + Lc_{j+s}^T y_{j+s}
//   backward (node i = s-1+2su):    x_i -= Lc_i x_{i-s} + R_i^T x_{i+s}
template <typename T, bool BWD>
__global__ __launch_bounds__(kThreads) void k_update_large(const T* __restrict__ E, T* __restrict__ x,
                                                          const uint8_t* __restrict__ mask, int N, int n, int m,
                                                          int nE, int s, int off) {
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int NB = (n + 31) / 32;
  const int groups = (NB + 3) / 4;
  const int rb = 4 * (blockIdx.y % groups) + wave_id();
  const int c0 = 32 * (blockIdx.y / groups);
  if (rb >= NB || c0 >= m) return;  // whole wavefronts: no barrier below
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  const T* Eb = E + (size_t)b * nE * nn;
  T* xb = x + (size_t)b * N * nm;
  int j, ja = -1, jc = -1;
  const T* Aop = nullptr;  // times x_ja, rows of the target
  const T* Cop = nullptr;  // transposed, times x_jc
  if (!BWD) {
    j = 2 * s * (blockIdx.x + 1) - 1;
    ja = j - s;
    Aop = Eb + (size_t)(off + (j - s) / s) * nn;  // R of node j-s: block (j, j-s)
    if (j + s < N) {
      jc = j + s;
      Cop = Eb + (size_t)(off + j / s) * nn;  // Lc of node j+s: block (j+s, j)
    }
  } else {
    j = s - 1 + 2 * s * blockIdx.x;
    if (j >= s) {
      ja = j - s;
      Aop = Eb + (size_t)(off + (j - s) / s) * nn;  // Lc_j: block (j, j-s)
    }
    if (j + s < N) {
      jc = j + s;
      Cop = Eb + (size_t)(off + j / s) * nn;  // R_j: block (j+s, j)
    }
    if (ja < 0 && jc < 0) return;
  }
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* X0 = reinterpret_cast<T*>(smem_raw) + 2 * wave_id() * kTileElems;
  T* X1 = X0 + kTileElems;
  typename Mfma<T>::acc_t acc[2][2];
  acc_clear<T>(acc);
  for (int kb = 0; kb < NB; ++kb) {
    if (Aop) {  // A[rb, kb] x_ja[kb, c0..]
      wtile_load(X0, Aop, n, n, n, 32 * rb, 32 * kb, false);
      wtile_load(X1, xb + (size_t)ja * nm, n, m, m, 32 * kb, c0, false);
      wave_gemm<T>(acc, X0, kLP, 1, X1, kLP, 1, 32);
    }
    if (Cop) {  // C[kb, rb]^T x_jc[kb, c0..]
      wtile_load(X0, Cop, n, n, n, 32 * kb, 32 * rb, false);
      wtile_load(X1, xb + (size_t)jc * nm, n, m, m, 32 * kb, c0, false);
      wave_gemm<T>(acc, X0, 1, kLP, X1, kLP, 1, 32);
    }
  }
  acc_to_global<T>(acc, xb + (size_t)j * nm, n, m, m, 32 * rb, c0, 1);
}

constexpr size_t update_large_lds_elems() { return 8 * (size_t)kTileElems; }

template <typename T>
hipError_t set_large_attrs() {
  hipError_t e = hipFuncSetAttribute((const void*)k_potrf_large<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)potrf_large_lds<T>());
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_trtri_large<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)trtri_large_lds<T>());
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_trimul_large<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)trimul_large_lds<T>());
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_update_large<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(T) * update_large_lds_elems()));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_update_large<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(T) * update_large_lds_elems()));
  return e;
}

}  // namespace

hipError_t init_large_kernels() {
  hipError_t e = set_large_attrs<float>();
  return e == hipSuccess ? set_large_attrs<double>() : e;
}

template <typename T>
hipError_t potrf_large(T* L, const uint8_t* mask, int B, int N, int n, int lidx, hipStream_t st) {
  const Level lv = level_info(N, lidx);
  if (lv.dim <= 0 || B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_potrf_large<T>, dim3(lv.dim, B), dim3(kThreads), potrf_large_lds<T>(), st, L, mask, N, n,
                     lv.s);
  return hipGetLastError();
}

template <typename T>
hipError_t trsm_large(const T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, hipStream_t st) {
  const Level lv = level_info(N, lidx);
  if (lv.dim <= 0 || B <= 0) return hipSuccess;
  const int NB = (n + 31) / 32;
  const int nE = (int)off_diag_storage_len(N);
  hipLaunchKernelGGL(k_trimul_large<T>, dim3(lv.dim, 2 * NB, B), dim3(kThreads), trimul_large_lds<T>(), st, L, E,
                     (T*)nullptr, mask, N, n, 1, nE, lv.s, lv.off, (int)kTrsm);
  return hipGetLastError();
}

template <typename T>
hipError_t apply_large(const T* L, T* x, const uint8_t* mask, int B, int N, int n, int m, int lidx, bool bwd,
                       hipStream_t st) {
  const Level lv = level_info(N, lidx);
  if (lv.dim <= 0 || B <= 0) return hipSuccess;
  const int mch = (m + 31) / 32;
  hipLaunchKernelGGL(k_trimul_large<T>, dim3(lv.dim, mch, B), dim3(kThreads), trimul_large_lds<T>(), st, L,
                     (T*)nullptr, x, mask, N, n, m, 0, lv.s, 0, bwd ? (int)kApplyBwd : (int)kApplyFwd);
  return hipGetLastError();
}

template <typename T>
hipError_t update_large(const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, int lidx, bool bwd,
                        hipStream_t st) {
  const Level lv = level_info(N, lidx);
  const int items = bwd ? lv.dim : N / (2 * lv.s);
  if (items <= 0 || B <= 0) return hipSuccess;
  const int NB = (n + 31) / 32;
  const int nE = (int)off_diag_storage_len(N);
  const int grid_y = ((NB + 3) / 4) * ((m + 31) / 32);
  const size_t lds = sizeof(T) * update_large_lds_elems();
  if (bwd)
    hipLaunchKernelGGL((k_update_large<T, true>), dim3(items, grid_y, B), dim3(kThreads), lds, st, E, x, mask, N, n,
                       m, nE, lv.s, lv.off);
  else
    hipLaunchKernelGGL((k_update_large<T, false>), dim3(items, grid_y, B), dim3(kThreads), lds, st, E, x, mask, N, n,
                       m, nE, lv.s, lv.off);
  return hipGetLastError();
}

template <typename T>
hipError_t trtri_large(const T* L, T* W, const uint8_t* mask, int B, int N, int n, hipStream_t st) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_trtri_large<T>, dim3(N, B), dim3(kThreads), trtri_large_lds<T>(), st, L, W, mask, N, n);
  return hipGetLastError();
}

template hipError_t potrf_large<float>(float*, const uint8_t*, int, int, int, int, hipStream_t);
template hipError_t potrf_large<double>(double*, const uint8_t*, int, int, int, int, hipStream_t);
template hipError_t trsm_large<float>(const float*, float*, const uint8_t*, int, int, int, int, hipStream_t);
template hipError_t trsm_large<double>(const double*, double*, const uint8_t*, int, int, int, int, hipStream_t);
template hipError_t apply_large<float>(const float*, float*, const uint8_t*, int, int, int, int, int, bool,
                                       hipStream_t);
template hipError_t apply_large<double>(const double*, double*, const uint8_t*, int, int, int, int, int, bool,
                                        hipStream_t);
template hipError_t update_large<float>(const float*, float*, const uint8_t*, int, int, int, int, int, bool,
                                        hipStream_t);
template hipError_t update_large<double>(const double*, double*, const uint8_t*, int, int, int, int, int, bool,
                                         hipStream_t);
template hipError_t trtri_large<float>(const float*, float*, const uint8_t*, int, int, int, hipStream_t);
template hipError_t trtri_large<double>(const double*, double*, const uint8_t*, int, int, int, hipStream_t);

}  // namespace socu
