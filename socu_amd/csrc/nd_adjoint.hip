// Adjoint of x = Psi^{-1} b for the block-tridiagonal Psi of the ND solver
// (diagonal blocks sym(L_i) built from the lower triangle of L_i, coupling
// blocks E_i at (i+1, i) and E_i^T at (i, i+1)).  With bbar = Psi^{-1} xbar
// (the transpose solve, made with the same factor since Psi is symmetric)
// the gradient of the full matrix is G = -bbar x^T; restricted to the
// parameters it is
//   dL_i[r][c] = -(bbar_i[r] . x_i[c] + x_i[r] . bbar_i[c])   r > c
//   dL_i[r][r] = -(bbar_i[r] . x_i[r]),   dL_i[r][c] = 0       r < c
//   dE_e[r][c] = -(bbar_{e+1}[r] . x_e[c] + x_{e+1}[r] . bbar_e[c])   e < N-1
//   dE_e = 0 on the fill-in region (e >= N-1)
// where "." sums over the m right-hand sides.  One workgroup computes one
// 32 x 32 tile of one output block, staging the four vector slices it needs
// through LDS in chunks of CH right-hand sides (8 for m <= 8: the staging
// arrays then take 9 KiB instead of 34 KiB, and twice as many workgroups fit
// on a CU); every output element is written exactly once.
#include "../../include/socu_hip.h"
#include "nd_common.h"
#include "nd_internal.h"

namespace socu {

namespace {

constexpr int kAdjTile = 32;

template <typename T, int CH>
__global__ __launch_bounds__(256) void k_adjoint_scatter(const T* __restrict__ bbar, const T* __restrict__ x,
                                                         T* __restrict__ dL, T* __restrict__ dE,
                                                         const uint8_t* __restrict__ mask, int N, int n, int m,
                                                         long long nE, int tiles1d) {
  __shared__ T P[kAdjTile][CH + 1];  // rows side: first product
  __shared__ T Q[kAdjTile][CH + 1];  // cols side: first product
  __shared__ T R[kAdjTile][CH + 1];  // rows side: second product
  __shared__ T S[kAdjTile][CH + 1];  // cols side: second product
  const int b = blockIdx.y;
  const int tiles = tiles1d * tiles1d;
  const long long blk = blockIdx.x / tiles;
  const int tile = blockIdx.x - (int)(blk * tiles);
  const int r0 = (tile / tiles1d) * kAdjTile, c0 = (tile % tiles1d) * kAdjTile;
  const size_t nn = (size_t)n * n;
  const bool isL = blk < N;
  const long long e = blk - N;  // E block index when !isL
  T* out = isL ? dL + ((size_t)b * N + blk) * nn : dE + ((size_t)b * nE + e) * nn;
  const int tid = threadIdx.x;
  const int tr = tid / 8, tc8 = tid % 8;  // thread owns row r0+tr, cols c0+tc8+8q
  if (masked_out(mask, b)) return;  // masked systems keep their outputs
  const bool fill = !isL && e >= N - 1;
  T acc[4] = {T(0), T(0), T(0), T(0)};
  if (!fill) {
    const size_t nm = (size_t)n * m;
    const T* bb = bbar + (size_t)b * N * nm;
    const T* xx = x + (size_t)b * N * nm;
    // (rows, cols) sources: dL_i: (bbar_i, x_i) + (x_i, bbar_i);
    // dE_e: (bbar_{e+1}, x_e) + (x_{e+1}, bbar_e)
    const size_t ra = isL ? (size_t)blk : (size_t)(e + 1), ca = isL ? (size_t)blk : (size_t)e;
    const T* pr = bb + ra * nm;
    const T* qc = xx + ca * nm;
    const T* rr = xx + ra * nm;
    const T* sc = bb + ca * nm;
    for (int k0 = 0; k0 < m; k0 += CH) {
      const int kc = min(CH, m - k0);
      for (int idx = tid; idx < kAdjTile * CH; idx += 256) {
        const int i = idx / CH, k = idx % CH;
        const bool okr = r0 + i < n && k < kc, okc = c0 + i < n && k < kc;
        P[i][k] = okr ? pr[(size_t)(r0 + i) * m + k0 + k] : T(0);
        R[i][k] = okr ? rr[(size_t)(r0 + i) * m + k0 + k] : T(0);
        Q[i][k] = okc ? qc[(size_t)(c0 + i) * m + k0 + k] : T(0);
        S[i][k] = okc ? sc[(size_t)(c0 + i) * m + k0 + k] : T(0);
      }
      __syncthreads();
      for (int k = 0; k < kc; ++k) {
        const T p = P[tr][k], q = R[tr][k];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] += p * Q[tc8 + 8 * j][k] + q * S[tc8 + 8 * j][k];
      }
      __syncthreads();
    }
  }
  const int r = r0 + tr;
  if (r < n) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = c0 + tc8 + 8 * j;
      if (c >= n) continue;
      T v = fill ? T(0) : -acc[j];
      if (isL) v = (r > c) ? v : ((r == c) ? T(0.5) * v : T(0));
      out[(size_t)r * n + c] = v;
    }
  }
}

}  // namespace

hipError_t adjoint_scatter(const void* bbar, const void* x, void* dL, void* dE, const uint8_t* mask, int B, int N,
                           int n, int m, int dtype, hipStream_t st) {
  const long long nE = off_diag_storage_len(N);
  const int tiles1d = (n + kAdjTile - 1) / kAdjTile;
  const long long gx = (N + nE) * tiles1d * tiles1d;
  const bool narrow = m <= 8;
  if (dtype == SOCU_DTYPE_F64) {
    auto k = narrow ? k_adjoint_scatter<double, 8> : k_adjoint_scatter<double, 32>;
    hipLaunchKernelGGL(k, dim3((unsigned)gx, B), dim3(256), 0, st, static_cast<const double*>(bbar),
                       static_cast<const double*>(x), static_cast<double*>(dL), static_cast<double*>(dE), mask, N, n,
                       m, nE, tiles1d);
  } else {
    auto k = narrow ? k_adjoint_scatter<float, 8> : k_adjoint_scatter<float, 32>;
    hipLaunchKernelGGL(k, dim3((unsigned)gx, B), dim3(256), 0, st, static_cast<const float*>(bbar),
                       static_cast<const float*>(x), static_cast<float*>(dL), static_cast<float*>(dE), mask, N, n, m,
                       nE, tiles1d);
  }
  return hipGetLastError();
}

}  // namespace socu
