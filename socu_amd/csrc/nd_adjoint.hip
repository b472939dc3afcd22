// Adjoint scatters of the ND solver (no counterpart in the reference, whose
// launches are not differentiable).
//
// Mode kPsi -- x = Psi^{-1} b for the block-tridiagonal Psi (diagonal blocks
// sym(L_i) built from the lower triangle of L_i, coupling blocks E_i at
// (i+1, i) and E_i^T at (i, i+1)).  With bbar = Psi^{-1} xbar (the transpose
// solve, made with the same factor since Psi is symmetric) the gradient of the
// full matrix is G = -bbar x^T; restricted to the parameters it is
//   dL_i[r][c] = -(bbar_i[r] . x_i[c] + x_i[r] . bbar_i[c])   r > c
//   dL_i[r][r] = -(bbar_i[r] . x_i[r]),   dL_i[r][c] = 0       r < c
//   dE_e[r][c] = -(bbar_{e+1}[r] . x_e[c] + x_{e+1}[r] . bbar_e[c])   e < N-1
//   dE_e = 0 on the fill-in region (e >= N-1)
// where "." sums over the m right-hand sides.
//
// Mode kFactor -- the gradient with respect to the entries of the ND factor
// Lp (PΨP^T = Lp Lp^T) as stored by the factor: G = -(u v^T + p q^T) on the
// sparsity pattern of Lp, with u, v, p, q given in the solver's unpermuted
// vector layout (p, q optional).  Diagonal block i: dL_i = tril(G_ii) (the
// diagonal not halved: it holds factor entries, not a symmetric pair).  E
// entry k of level segment l couples a = (k+1)s-1 and b = (k+2)s-1 (s = 2^l)
// and holds Lp[b][a] for even k, Lp[a][b]^T for odd k
// (block_tridiag_solver.py:92-145 index maps), so
//   even k: dE = -(u_b v_a^T + p_b q_a^T),   odd k: dE = -(v_b u_a^T + q_b p_a^T).
// Every E index is an entry of Lp (each is written by the factor), and the
// strict upper triangle of each L_i gets zeros.  For y = Lp^{-1} P b the
// gradient is G = -bbar y^T with bbar = Lp^{-T} ybar; for x = P^T Lp^{-T} y it
// is G = -x w^T with w = Lp^{-1} xbar; a solve adds both.
//
// One workgroup computes one 32 x 32 tile of one output block, staging the
// four vector slices it needs through LDS in chunks of CH right-hand sides (8
// for m <= 8: the staging arrays then take 9 KiB instead of 34 KiB, and twice
// as many workgroups fit on a CU); every output element is written exactly
// once.
#include "../../include/socu_hip.h"
#include "nd_common.h"
#include "nd_internal.h"

namespace socu {

namespace {

constexpr int kAdjTile = 32;
enum { kPsi = 0, kFactor = 1 };

// Vector block of the (rows, cols) sides of the two products.  kPsi:
// (u, v, p, q) = (bbar, x, x, bbar).
template <typename T>
struct AdjSrc {
  const T* P;  // rows side, first product
  const T* Q;  // cols side, first product
  const T* R;  // rows side, second product (NULL: none)
  const T* S;  // cols side, second product
};

template <typename T, int CH, int MODE>
__global__ __launch_bounds__(256) void k_adjoint_scatter(const T* __restrict__ u, const T* __restrict__ v,
                                                         const T* __restrict__ p, const T* __restrict__ q,
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
  const int tc = tid & 31, tr = tid >> 5;  // thread owns column c0+tc, rows r0+tr+8q: each store
                                          // instruction writes two whole 32-wide rows
  if (masked_out(mask, b)) return;  // masked systems keep their outputs
  const bool fill = MODE == kPsi && !isL && e >= N - 1;
  T acc[4] = {T(0), T(0), T(0), T(0)};
  if (!fill) {
    const size_t nm = (size_t)n * m;
    const size_t sys = (size_t)b * N * nm;
    // row / column vector blocks and which inputs feed which side
    size_t ra, ca;
    bool swap = false;  // odd factor entries: rows from (v, q), cols from (u, p)
    if (isL) {
      ra = ca = (size_t)blk;
    } else if (MODE == kPsi) {
      ra = (size_t)(e + 1);
      ca = (size_t)e;
    } else {
      long long off = 0;
      int l = 0;
      while (off + (N >> l) - 1 <= e) {
        off += (N >> l) - 1;
        ++l;
      }
      const long long k = e - off;
      const long long s = 1ll << l;
      ca = (size_t)((k + 1) * s - 1);  // a
      ra = (size_t)((k + 2) * s - 1);  // b
      swap = (k & 1) != 0;
    }
    AdjSrc<T> src;
    if (MODE == kPsi) {
      src = {u + sys + ra * nm, v + sys + ca * nm, p + sys + ra * nm, q + sys + ca * nm};
    } else if (!swap) {
      src = {u + sys + ra * nm, v + sys + ca * nm, p ? p + sys + ra * nm : nullptr, p ? q + sys + ca * nm : nullptr};
    } else {
      src = {v + sys + ra * nm, u + sys + ca * nm, p ? q + sys + ra * nm : nullptr, p ? p + sys + ca * nm : nullptr};
    }
    const bool two = src.R != nullptr;
    for (int k0 = 0; k0 < m; k0 += CH) {
      const int kc = min(CH, m - k0);
      // unconditional loads from clamped addresses, all four in flight
      // before any is used, then the masking (a select on each load let the
      // compiler give every load a branch and a wait of its own)
      const T* Rp = two ? src.R : src.P;
      const T* Sp = two ? src.S : src.Q;
      for (int idx = tid; idx < kAdjTile * CH; idx += 256) {
        const int i = idx / CH, k = idx % CH;
        const int kk = min(k, kc - 1);
        const size_t ro = (size_t)min(r0 + i, n - 1) * m + k0 + kk, co = (size_t)min(c0 + i, n - 1) * m + k0 + kk;
        T pv = src.P[ro], qv = src.Q[co], rv = Rp[ro], sv = Sp[co];
        pin_reg(pv);
        pin_reg(qv);
        pin_reg(rv);
        pin_reg(sv);
        const bool okr = r0 + i < n && k < kc, okc = c0 + i < n && k < kc;
        P[i][k] = okr ? pv : T(0);
        Q[i][k] = okc ? qv : T(0);
        R[i][k] = okr && two ? rv : T(0);
        S[i][k] = okc && two ? sv : T(0);
      }
      __syncthreads();
      for (int k = 0; k < kc; ++k) {
        const T qv = Q[tc][k], sv = S[tc][k];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] += P[tr + 8 * j][k] * qv + R[tr + 8 * j][k] * sv;
      }
      __syncthreads();
    }
  }
  const int c = c0 + tc;
  if (c < n) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = r0 + tr + 8 * j;
      if (r >= n) continue;
      T val = fill ? T(0) : -acc[j];
      if (isL) val = (r > c) ? val : ((r == c) ? (MODE == kPsi ? T(0.5) * val : val) : T(0));
      out[(size_t)r * n + c] = val;
    }
  }
}

template <typename T, int MODE>
void launch_adjoint(const void* u, const void* v, const void* p, const void* q, void* dL, void* dE,
                    const uint8_t* mask, int B, int N, int n, int m, hipStream_t st) {
  const long long nE = off_diag_storage_len(N);
  const int tiles1d = (n + kAdjTile - 1) / kAdjTile;
  const long long gx = (N + nE) * tiles1d * tiles1d;
  auto k = m <= 8 ? k_adjoint_scatter<T, 8, MODE> : k_adjoint_scatter<T, 32, MODE>;
  hipLaunchKernelGGL(k, dim3((unsigned)gx, B), dim3(256), 0, st, static_cast<const T*>(u), static_cast<const T*>(v),
                     static_cast<const T*>(p), static_cast<const T*>(q), static_cast<T*>(dL), static_cast<T*>(dE),
                     mask, N, n, m, nE, tiles1d);
}

}  // namespace

hipError_t adjoint_scatter(const void* bbar, const void* x, void* dL, void* dE, const uint8_t* mask, int B, int N,
                           int n, int m, int dtype, hipStream_t st) {
  if (dtype == SOCU_DTYPE_F64) launch_adjoint<double, kPsi>(bbar, x, x, bbar, dL, dE, mask, B, N, n, m, st);
  else launch_adjoint<float, kPsi>(bbar, x, x, bbar, dL, dE, mask, B, N, n, m, st);
  return hipGetLastError();
}

hipError_t factor_adjoint_scatter(const void* u, const void* v, const void* p, const void* q, void* dL, void* dE,
                                  const uint8_t* mask, int B, int N, int n, int m, int dtype, hipStream_t st) {
  if (dtype == SOCU_DTYPE_F64) launch_adjoint<double, kFactor>(u, v, p, q, dL, dE, mask, B, N, n, m, st);
  else launch_adjoint<float, kFactor>(u, v, p, q, dL, dE, mask, B, N, n, m, st);
  return hipGetLastError();
}

}  // namespace socu
