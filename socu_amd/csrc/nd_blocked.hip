// Blocked ND kernels for block size n <= 128 on gfx950: blocks are padded to
// NP = 32*NB (NB = ceil(n/32)) inside LDS/registers only; HBM accesses stay
// within the n x n blocks.  One kernel launch per phase per level; no atomics.
//
//   k_potrf_blk  one workgroup per node:   L_i <- chol(L_i), blocked
//                right-looking over 32x32 tiles (reference: potrf_blocked,
//                socu/blas/potrf.py:97-119).  Each diagonal tile is factored
//                and inverted by one wavefront (tile_chol_inv); the panel and
//                trailing updates run on MFMA.  The transposed tile inverses
//                are left in the strict upper triangle of the diagonal tiles
//                (the upper triangle of L is not part of the factor).
//   k_trsm_blk   one workgroup per (node, 32-wide chunk of R rows / Lc cols):
//                R <- R L^{-T}, Lc <- L^{-1} Lc by blocked substitution with
//                the tile inverses (trsm.py:94-120, :202-228).
//   k_schur_blk  one workgroup per 64x64 output tile, MFMA, owner computes:
//                L_j -= R R^T + Lc^T Lc,  F -= R Lc (syrk.py, gemm.py).
//   k_fwd_blk / k_bwd_blk  forward step / backward level (see nd_solve.hip
//                for the step structure): coalesced block-times-panel
//                products and blocked triangular solves with one-wave tile
//                solves.
#include "nd_common.h"
#include "nd_internal.h"
#include "nd_tile.h"

namespace socu {

// ---------------------------------------------------------------------------
// potrf
// ---------------------------------------------------------------------------
// NF != 0 instantiates the kernels for a fixed block size n == NF (a multiple
// of 32), so that all index arithmetic is compile-time; NF == 0 is the
// general (padded) version.
template <typename T, int NB, int NF>
__global__ __launch_bounds__(kThreads) void k_potrf_blk(T* __restrict__ L, const uint8_t* __restrict__ mask, int N,
                                                       int n_, int s) {
  const int n = NF ? NF : n_;
  constexpr int NP = 32 * NB;
  constexpr int P = NP + 1;
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  T* A = L + ((size_t)b * N + i) * (size_t)n * n;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);
  T* dinv = S + NP * P;
  T* col = dinv + 32;
  const int tid = threadIdx.x, wave = wave_id(), l = lane_id();

  // The block is read as one contiguous n*n run (coalesced, unconditional
  // loads from clamped addresses) and scattered into the padded LDS square.
  constexpr int PER = NP * NP / kThreads;
  const int nn = n * n;
  T v[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * kThreads;
    v[e] = A[min(idx, nn - 1)];
  }
  for (int idx = tid; idx < NP * NP; idx += kThreads) {
    const int r = idx / NP, c = idx - r * NP;
    S[r * P + c] = (r == c) ? T(1) : T(0);
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * kThreads;
    const int r = idx / n, c = idx - r * n;
    if (idx < nn) S[r * P + c] = (c <= r) ? v[e] : T(0);
  }
  __syncthreads();

  for (int kb = 0; kb < NB; ++kb) {
    T* D = S + 32 * kb * P + 32 * kb;
    if (wave == 0) tile_chol_inv<T>(D, P, dinv, col);
    __syncthreads();
    // panel: tile (ti, kb) <- tile (ti, kb) Linv^T
    for (int ti = kb + 1 + wave; ti < NB; ti += kThreads / 64) {
      T* X = S + 32 * ti * P + 32 * kb;
      typename Mfma<T>::acc_t acc[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
      const int li = l & 15, kq = l >> 4;
#pragma unroll
      for (int k = 0; k < 32; k += 4) {
        const int kk = k + kq;
        const T a0 = X[li * P + kk], a1 = X[(li + 16) * P + kk];
        // B(k, c) = Linv[c][k]: strict upper of D holds it for c > k
        const int c0 = li, c1 = li + 16;
        const T b0 = (c0 > kk) ? D[kk * P + c0] : ((c0 == kk) ? dinv[kk] : T(0));
        const T b1 = (c1 > kk) ? D[kk * P + c1] : ((c1 == kk) ? dinv[kk] : T(0));
        acc[0][0] = Mfma<T>::mma(a0, b0, acc[0][0]);
        acc[0][1] = Mfma<T>::mma(a0, b1, acc[0][1]);
        acc[1][0] = Mfma<T>::mma(a1, b0, acc[1][0]);
        acc[1][1] = Mfma<T>::mma(a1, b1, acc[1][1]);
      }
      acc_foreach<T>(acc, [&](int r, int c, T v) { X[r * P + c] = v; });
    }
    __syncthreads();
    // trailing update of the lower tiles (ti, tj), kb < tj <= ti
    const int m = NB - 1 - kb;
    const int cnt = m * (m + 1) / 2;
    for (int t = wave; t < cnt; t += kThreads / 64) {
      int ti = 0, q = t;
      while (q > ti) { q -= ti + 1; ++ti; }
      const int tj = q;
      const int gi = kb + 1 + ti, gj = kb + 1 + tj;
      typename Mfma<T>::acc_t acc[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
      wave_gemm<T>(acc, S + 32 * gi * P + 32 * kb, P, 1, S + 32 * gj * P + 32 * kb, 1, P, 32);
      T* C = S + 32 * gi * P + 32 * gj;
      acc_foreach<T>(acc, [&](int r, int c, T v) { C[r * P + c] -= v; });
    }
    __syncthreads();
  }

  // Whole block written back as one contiguous run: tril(L) is the factor;
  // the strict upper triangle (not part of the factor) carries the diagonal
  // tiles' transposed inverses and stale values elsewhere.
  for (int idx = tid; idx < nn; idx += kThreads) {
    const int r = idx / n, c = idx - r * n;
    A[idx] = S[r * P + c];
  }
}

// ---------------------------------------------------------------------------
// trsm
// ---------------------------------------------------------------------------
template <typename T, int NB, int NF>
__global__ __launch_bounds__(kThreads) void k_trsm_blk(const T* __restrict__ L, T* __restrict__ E,
                                                      const uint8_t* __restrict__ mask, int N, int n_, int nE, int s,
                                                      int off) {
  const int n = NF ? NF : n_;
  constexpr int NP = 32 * NB;
  constexpr int TP = 33;  // tile pitch
  constexpr int NT = NB * (NB + 1) / 2;
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  const bool isR = (int)blockIdx.y < NB;
  const int c0 = (isR ? blockIdx.y : blockIdx.y - NB) * 32;
  if (isR && !(i + s < N)) return;
  if (!isR && !(i >= s)) return;
  if (c0 >= n) return;
  const int w = min(32, n - c0);

  const T* A = L + ((size_t)b * N + i) * (size_t)n * n;
  T* X = E + ((size_t)b * nE + off + (isR ? i / s : (i - s) / s)) * (size_t)n * n;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Lt = reinterpret_cast<T*>(smem_raw);  // NT tiles of 32 x TP
  T* dinv = Lt + NT * 32 * TP;             // NP
  T* Z = dinv + NP;                        // NP x TP
  const int tid = threadIdx.x, wave = wave_id(), l = lane_id();

  // Issue every load (tiles of L_i, pivots, the panel) before storing any.
  constexpr int PT = 32 * 32 / kThreads;  // elements per thread per tile
  T lv[NT][PT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    int ti = 0, q = t;
    while (q > ti) { q -= ti + 1; ++ti; }
    const int tj = q;
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = tid + e * kThreads;
      const int R = 32 * ti + (idx >> 5), C = 32 * tj + (idx & 31);
      const T g = A[(size_t)min(R, n - 1) * n + min(C, n - 1)];
      lv[t][e] = (R < n && C < n) ? g : ((R == C) ? T(1) : T(0));
    }
  }
  constexpr int PZ = 32 * NP / kThreads;
  T zv[PZ];
  if (isR) {
#pragma unroll
    for (int e = 0; e < PZ; ++e) {
      const int idx = tid + e * kThreads;
      const int c = idx / NP, k = idx - c * NP;
      const T g = X[(size_t)(c0 + min(c, w - 1)) * n + min(k, n - 1)];
      zv[e] = (c < w && k < n) ? g : T(0);
    }
  } else {
#pragma unroll
    for (int e = 0; e < PZ; ++e) {
      const int idx = tid + e * kThreads;
      const int k = idx >> 5, c = idx & 31;
      const T g = X[(size_t)min(k, n - 1) * n + c0 + min(c, w - 1)];
      zv[e] = (c < w && k < n) ? g : T(0);
    }
  }
  for (int k = tid; k < NP; k += kThreads) dinv[k] = (k < n) ? T(1) / A[(size_t)k * n + k] : T(1);
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = tid + e * kThreads;
      Lt[t * 32 * TP + (idx >> 5) * TP + (idx & 31)] = lv[t][e];
    }
  if (isR) {
#pragma unroll
    for (int e = 0; e < PZ; ++e) {
      const int idx = tid + e * kThreads;
      const int c = idx / NP, k = idx - c * NP;
      Z[k * TP + c] = zv[e];
    }
  } else {
#pragma unroll
    for (int e = 0; e < PZ; ++e) {
      const int idx = tid + e * kThreads;
      Z[(idx >> 5) * TP + (idx & 31)] = zv[e];
    }
  }
  __syncthreads();

  const int qi = wave >> 1, qj = wave & 1;
  const int li = l & 15, kq = l >> 4;
  for (int kb = 0; kb < NB; ++kb) {
    if (kb > 0) {
      typename Mfma<T>::acc_t acc = acc_zero<T>();
      for (int j = 0; j < kb; ++j) {
        const T* At = Lt + (kb * (kb + 1) / 2 + j) * 32 * TP + 16 * qi * TP;
        wave_gemm16<T>(acc, At, TP, 1, Z + 32 * j * TP + 16 * qj, TP, 1, 32);
      }
      T* Zq = Z + (32 * kb + 16 * qi) * TP + 16 * qj;
      acc16_foreach<T>(acc, [&](int r, int c, T v) { Zq[r * TP + c] -= v; });
      __syncthreads();
    }
    // Z[kb] <- Linv_kk Z[kb]; A(r,k) = Linv[r][k] = D[k][r] (r > k), dinv (r == k)
    const T* D = Lt + (kb * (kb + 1) / 2 + kb) * 32 * TP;
    typename Mfma<T>::acc_t acc = acc_zero<T>();
#pragma unroll
    for (int k = 0; k < 32; k += 4) {
      const int kk = k + kq;
      const int r = 16 * qi + li;
      const T a = (r > kk) ? D[kk * TP + r] : ((r == kk) ? dinv[32 * kb + kk] : T(0));
      const T bv = Z[(32 * kb + kk) * TP + 16 * qj + li];
      acc = Mfma<T>::mma(a, bv, acc);
    }
    __syncthreads();
    T* Zq = Z + (32 * kb + 16 * qi) * TP + 16 * qj;
    acc16_foreach<T>(acc, [&](int r, int c, T v) { Zq[r * TP + c] = v; });
    __syncthreads();
  }

  if (isR) {
    for (int idx = tid; idx < w * n; idx += kThreads) {
      const int c = idx / n, k = idx - c * n;
      X[(size_t)(c0 + c) * n + k] = Z[k * TP + c];
    }
  } else {
    for (int idx = tid; idx < n * w; idx += kThreads) {
      const int k = idx / w, c = idx - k * w;
      X[(size_t)k * n + c0 + c] = Z[k * TP + c];
    }
  }
}

// ---------------------------------------------------------------------------
// Schur-complement updates, 64x64 output tiles on MFMA
// ---------------------------------------------------------------------------
constexpr int kSTile = 64;
constexpr int kSK = 32;
constexpr int kSP = kSTile + 1;

enum { kABt = 0, kAtB = 1, kAB = 2 };

// One K-chunk of both operands of a 64x64 tile product, held in registers
// between the global loads and the LDS stores so that the next chunk's loads
// are in flight while the current chunk is multiplied.
// As[kk][a] = A(r0+a, k0+kk), Bs[kk][a] = B(k0+kk, c0+a), zero outside n.
constexpr int kSPer = kSTile * kSK / kThreads;  // elements per thread per operand

template <typename T>
struct Chunk64 {
  T a[kSPer];
  T b[kSPer];
};

template <typename T>
__device__ __forceinline__ void load_chunk64(Chunk64<T>& ch, const T* __restrict__ Am, const T* __restrict__ Bm,
                                             int form, int n, int r0, int c0, int k0) {
#pragma unroll
  for (int e = 0; e < kSPer; ++e) {
    const int idx = threadIdx.x + e * kThreads;
    const int a = idx / kSK, kk = idx - a * kSK, k = k0 + kk;                // k fastest: row reads
    const int kk2 = idx / kSTile, a2 = idx - kk2 * kSTile, k2 = k0 + kk2;  // a fastest: column reads
    // unconditional loads from clamped addresses, then select
    const int ra = min(r0 + a, n - 1), ca = min(c0 + a, n - 1), kc = min(k, n - 1);
    const int ra2 = min(r0 + a2, n - 1), ca2 = min(c0 + a2, n - 1), kc2 = min(k2, n - 1);
    const bool va = r0 + a < n && k < n, vb = c0 + a < n && k < n;
    const bool va2 = r0 + a2 < n && k2 < n, vb2 = c0 + a2 < n && k2 < n;
    if (form == kABt) {
      const T x = Am[(size_t)ra * n + kc], y = Bm[(size_t)ca * n + kc];
      ch.a[e] = va ? x : T(0);
      ch.b[e] = vb ? y : T(0);
    } else if (form == kAtB) {
      const T x = Am[(size_t)kc2 * n + ra2], y = Bm[(size_t)kc2 * n + ca2];
      ch.a[e] = va2 ? x : T(0);
      ch.b[e] = vb2 ? y : T(0);
    } else {
      const T x = Am[(size_t)ra * n + kc], y = Bm[(size_t)kc2 * n + ca2];
      ch.a[e] = va ? x : T(0);
      ch.b[e] = vb2 ? y : T(0);
    }
  }
}

template <typename T>
__device__ __forceinline__ void store_chunk64(const Chunk64<T>& ch, T* As, T* Bs, int form) {
#pragma unroll
  for (int e = 0; e < kSPer; ++e) {
    const int idx = threadIdx.x + e * kThreads;
    const int a = idx / kSK, kk = idx - a * kSK;
    const int kk2 = idx / kSTile, a2 = idx - kk2 * kSTile;
    if (form == kAtB) As[kk2 * kSP + a2] = ch.a[e];
    else As[kk * kSP + a] = ch.a[e];
    if (form == kABt) Bs[kk * kSP + a] = ch.b[e];
    else Bs[kk2 * kSP + a2] = ch.b[e];
  }
}

// Sequence of up to two products over K chunks: acc += sum_p A_p B_p.
template <typename T>
__device__ __forceinline__ void products64(typename Mfma<T>::acc_t acc[2][2], bool active, int np, const T* const* Am,
                                           const T* const* Bm, const int* form, int n, int r0, int c0, T* As, T* Bs) {
  const int wave = wave_id();
  const int wi = wave >> 1, wj = wave & 1;
  const int nk = (n + kSK - 1) / kSK;
  const int total = np * nk;
  Chunk64<T> ch;
  load_chunk64<T>(ch, Am[0], Bm[0], form[0], n, r0, c0, 0);
  for (int q = 0; q < total; ++q) {
    const int p = q / nk;
    __syncthreads();
    store_chunk64<T>(ch, As, Bs, form[p]);
    __syncthreads();
    if (q + 1 < total) {
      const int p1 = (q + 1) / nk, k1 = ((q + 1) - p1 * nk) * kSK;
      load_chunk64<T>(ch, Am[p1], Bm[p1], form[p1], n, r0, c0, k1);
    }
    if (active) wave_gemm<T>(acc, As + 32 * wi, 1, kSP, Bs + 32 * wj, kSP, 1, kSK);
  }
}

template <typename T, int NF>
__global__ __launch_bounds__(kThreads) void k_schur_blk(T* __restrict__ L, T* __restrict__ E,
                                                       const uint8_t* __restrict__ mask, int N, int n_, int nE, int s,
                                                       int off, int next_off, int nT, int ntile) {
  const int n = NF ? NF : n_;
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  __shared__ T As[kSK * kSP];
  __shared__ T Bs[kSK * kSP];
  const int ntri = ntile * (ntile + 1) / 2;
  int task = blockIdx.x;
  const size_t nn = (size_t)n * n;
  T* Eb = E + (size_t)b * nE * nn;
  const int wave = wave_id();
  const int wi = wave >> 1, wj = wave & 1;
  typename Mfma<T>::acc_t acc[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
  T* C;
  int tr, tc;
  bool active = true;
  if (task < nT * ntri) {
    const int u = task / ntri;
    int q = task - u * ntri;
    tr = 0;
    while (q > tr) { q -= tr + 1; ++tr; }
    tc = q;
    active = !(tr == tc && wi < wj);  // upper 32x32 quarter of a diagonal tile
    const int j = 2 * s * (u + 1) - 1;
    C = L + ((size_t)b * N + j) * nn;
    const T* R = Eb + (size_t)(off + (j - s) / s) * nn;
    const T* Lc = (j + s < N) ? Eb + (size_t)(off + j / s) * nn : R;
    const T* Am[2] = {R, Lc};
    const int form[2] = {kABt, kAtB};
    products64<T>(acc, active, (j + s < N) ? 2 : 1, Am, Am, form, n, tr * kSTile, tc * kSTile, As, Bs);
  } else {
    task -= nT * ntri;
    const int f = task / (ntile * ntile);
    const int q = task - f * ntile * ntile;
    tr = q / ntile;
    tc = q - tr * ntile;
    const int i = s - 1 + 2 * s * (f + 1);
    const T* R = Eb + (size_t)(off + i / s) * nn;
    const T* Lc = Eb + (size_t)(off + (i - s) / s) * nn;
    C = Eb + (size_t)(next_off + (i - s) / (2 * s)) * nn;
    const T* Am[1] = {R};
    const T* Bm[1] = {Lc};
    const int form[1] = {kAB};
    products64<T>(acc, active, 1, Am, Bm, form, n, tr * kSTile, tc * kSTile, As, Bs);
  }
  if (!active) return;
  const int rb = tr * kSTile + 32 * wi, cb = tc * kSTile + 32 * wj;
  acc_foreach<T>(acc, [&](int r, int c, T v) {
    const int rr = rb + r, cc = cb + c;
    if (rr < n && cc < n) C[(size_t)rr * n + cc] -= v;
  });
}

// ---------------------------------------------------------------------------
// solves
// ---------------------------------------------------------------------------
// Y (n x w panel in LDS, pitch YP) -= A Y1 with A an n x n row-major block
// in HBM: 16 lanes per row, each holding a contiguous slice of the row, then a
// 16-lane shuffle reduction per right-hand side.  A wave covers 4 rows per
// pass; every pass's loads are issued before the first is used.
template <typename T, int NB>
__device__ __forceinline__ void panel_sub_AY(T* Y, const T* __restrict__ A, const T* Y1, int n, int w, int YP) {
  constexpr int NP = 32 * NB;
  constexpr int KPL = NP / 16;                     // elements per lane
  constexpr int PASSES = NP / (4 * (kThreads / 64));  // 4 rows per wave per pass
  const int l = lane_id(), wave = wave_id();
  const int sub = l & 15, rsub = l >> 4;
  T a[PASSES][KPL];
#pragma unroll
  for (int p = 0; p < PASSES; ++p) {
    const int r = 4 * (wave + p * (kThreads / 64)) + rsub;
#pragma unroll
    for (int e = 0; e < KPL; ++e) {
      const int k = sub * KPL + e;
      const T g = A[(size_t)min(r, n - 1) * n + min(k, n - 1)];
      a[p][e] = (r < n && k < n) ? g : T(0);
    }
  }
#pragma unroll
  for (int p = 0; p < PASSES; ++p) {
    const int r = 4 * (wave + p * (kThreads / 64)) + rsub;
    for (int c = 0; c < w; ++c) {
      T acc = T(0);
#pragma unroll
      for (int e = 0; e < KPL; ++e) acc += a[p][e] * Y1[(sub * KPL + e) * YP + c];
#pragma unroll
      for (int o = 8; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 16);
      if (sub == 0 && r < n) Y[r * YP + c] -= acc;
    }
  }
}

// Y -= A^T Y1: thread (r, part) holds A[k][r] for its K part (reads coalesced
// in r); the parts subtract in a fixed order, one barrier apart.
template <typename T, int NB>
__device__ __forceinline__ void panel_sub_AtY(T* Y, const T* __restrict__ A, const T* Y1, int n, int w, int YP) {
  constexpr int NP = 32 * NB;
  constexpr int PARTS = kThreads / NP;
  constexpr int KP = (NP + PARTS - 1) / PARTS;
  const int t = threadIdx.x;
  const int r = t % NP, part = t / NP;
  const bool act = part < PARTS && r < n;
  T a[KP];
#pragma unroll
  for (int e = 0; e < KP; ++e) {
    const int k = part * KP + e;
    const T g = A[(size_t)min(k, n - 1) * n + min(r, n - 1)];
    a[e] = (act && k < n) ? g : T(0);
  }
  for (int c = 0; c < w; ++c) {
    T acc = T(0);
#pragma unroll
    for (int e = 0; e < KP; ++e) {
      const int k = part * KP + e;
      if (k < NP) acc += a[e] * Y1[k * YP + c];
    }
    for (int p = 0; p < PARTS; ++p) {
      if (act && part == p) Y[r * YP + c] -= acc;
      __syncthreads();
    }
  }
}

// Y <- L^{-1} Y (forward, lower) using tril(L) from HBM; blocked over 32-row
// tiles, the diagonal tile solved by wave 0 with register broadcasts.
template <typename T, int NB>
__device__ __forceinline__ void panel_lower_solve(T* Y, const T* __restrict__ Lb, int n, int w, int YP) {
  const int l = lane_id(), wave = wave_id();
  for (int kb = 0; kb < NB; ++kb) {
    const int base = 32 * kb;
    if (wave == 0) {
      const int r = l & 31, R = base + r;
      T a[32];
      T rdiag = T(1);
#pragma unroll
      for (int c = 0; c < 32; ++c) {
        const int C = base + c;
        const T g = Lb[(size_t)min(R, n - 1) * n + min(C, n - 1)];
        a[c] = (R < n && C < n) ? ((c <= r) ? g : T(0)) : ((R == C) ? T(1) : T(0));
        if (c == r) rdiag = T(1) / a[c];
      }
      for (int c = 0; c < w; ++c) {
        T y = (R < n) ? Y[R * YP + c] : T(0);
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          const T yj = bcast_lane(y, j) * bcast_lane(rdiag, j);
          y = (r == j) ? yj : ((r > j) ? y - a[j] * yj : y);
        }
        if (l < 32 && R < n) Y[R * YP + c] = y;
      }
    }
    __syncthreads();
    // rows below the tile: Y[r] -= L[r][base:base+32] Y[base:base+32]
    const int rows = n - base - 32;
    if (rows > 0) {
      for (int q = threadIdx.x; q < rows * w; q += kThreads) {
        const int c = q / rows, r = base + 32 + (q - c * rows);
        T lrow[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) {
          const T g = Lb[(size_t)r * n + min(base + k, n - 1)];
          lrow[k] = (base + k < n) ? g : T(0);
        }
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < 32; ++k) acc += lrow[k] * Y[(base + k) * YP + c];
        Y[r * YP + c] -= acc;
      }
    }
    __syncthreads();
  }
}

// Y <- L^{-T} Y (backward) using tril(L) from HBM.
template <typename T, int NB>
__device__ __forceinline__ void panel_upper_solve_T(T* Y, const T* __restrict__ Lb, int n, int w, int YP) {
  const int l = lane_id(), wave = wave_id();
  for (int kb = NB - 1; kb >= 0; --kb) {
    const int base = 32 * kb;
    if (wave == 0) {
      // lane q holds column q of the diagonal tile: bcol[j] = L[base+j][base+q]
      const int q = l & 31, Q = base + q;
      T bcol[32];
      T rdiag = T(1);
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int R = base + j;
        const T g = Lb[(size_t)min(R, n - 1) * n + min(Q, n - 1)];
        bcol[j] = (R < n && Q < n) ? ((j >= q) ? g : T(0)) : ((R == Q) ? T(1) : T(0));
        if (j == q) rdiag = T(1) / bcol[j];
      }
      for (int c = 0; c < w; ++c) {
        T y = (Q < n) ? Y[Q * YP + c] : T(0);
#pragma unroll
        for (int j = 31; j >= 0; --j) {
          const T yj = bcast_lane(y, j) * bcast_lane(rdiag, j);
          y = (q == j) ? yj : ((q < j) ? y - bcol[j] * yj : y);
        }
        if (l < 32 && Q < n) Y[Q * YP + c] = y;
      }
    }
    __syncthreads();
    // rows above the tile: Y[r] -= L[base:base+32][r]^T Y[base:base+32]
    if (base > 0) {
      for (int qq = threadIdx.x; qq < base * w; qq += kThreads) {
        const int c = qq / base, r = qq - c * base;
        T lcol[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) {
          const T g = Lb[(size_t)min(base + k, n - 1) * n + r];
          lcol[k] = (base + k < n) ? g : T(0);
        }
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < 32; ++k) acc += lcol[k] * Y[(base + k) * YP + c];
        Y[r * YP + c] -= acc;
      }
    }
    __syncthreads();
  }
}

template <typename T>
__device__ __forceinline__ void load_y(T* Y, const T* __restrict__ X, int n, int m, int c0, int w, int YP, int NP) {
  for (int q = threadIdx.x; q < NP * w; q += kThreads) {
    const int r = q / w, c = q - r * w;
    Y[r * YP + c] = (r < n) ? X[(size_t)r * m + c0 + c] : T(0);
  }
}

template <typename T>
__device__ __forceinline__ void store_y(T* __restrict__ X, const T* Y, int n, int m, int c0, int w, int YP) {
  for (int q = threadIdx.x; q < n * w; q += kThreads) {
    const int r = q / w, c = q - r * w;
    X[(size_t)r * m + c0 + c] = Y[r * YP + c];
  }
}

template <typename T, int NB, int NF>
__global__ __launch_bounds__(kThreads) void k_fwd_blk(const T* __restrict__ L, const T* __restrict__ E,
                                                     T* __restrict__ x, const uint8_t* __restrict__ mask, int N, int n_,
                                                     int m, int nE, int s, int off, int first) {
  const int n = NF ? NF : n_;
  constexpr int NP = 32 * NB;
  constexpr int YP = kRhsChunk + 1;
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int u = blockIdx.x;
  const int c0 = blockIdx.y * kRhsChunk;
  const int w = min(kRhsChunk, m - c0);
  __shared__ T Y0[NP * YP];
  __shared__ T Y1[NP * YP];
  __shared__ T Y2[NP * YP];
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
  load_y(Y0, xb + (size_t)j * n * m, n, m, c0, w, YP, NP);
  if (!first) {
    const bool right = j + s < N;
    load_y(Y1, xb + (size_t)(j - s) * n * m, n, m, c0, w, YP, NP);
    if (right) load_y(Y2, xb + (size_t)(j + s) * n * m, n, m, c0, w, YP, NP);
    __syncthreads();
    // y_j -= R y_{j-s}, R = block (j, j-s)
    panel_sub_AY<T, NB>(Y0, Eb + (size_t)(off + (j - s) / s) * nn, Y1, n, w, YP);
    __syncthreads();
    // y_j -= Lc^T y_{j+s}, Lc = block (j+s, j)
    if (right) panel_sub_AtY<T, NB>(Y0, Eb + (size_t)(off + j / s) * nn, Y2, n, w, YP);
  }
  __syncthreads();
  if (do_solve) panel_lower_solve<T, NB>(Y0, L + ((size_t)b * N + j) * nn, n, w, YP);
  store_y(xb + (size_t)j * n * m, Y0, n, m, c0, w, YP);
}

template <typename T, int NB, int NF>
__global__ __launch_bounds__(kThreads) void k_bwd_blk(const T* __restrict__ L, const T* __restrict__ E,
                                                     T* __restrict__ x, const uint8_t* __restrict__ mask, int N, int n_,
                                                     int m, int nE, int s, int off) {
  const int n = NF ? NF : n_;
  constexpr int NP = 32 * NB;
  constexpr int YP = kRhsChunk + 1;
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  const int c0 = blockIdx.y * kRhsChunk;
  const int w = min(kRhsChunk, m - c0);
  __shared__ T Y0[NP * YP];
  __shared__ T Y1[NP * YP];
  __shared__ T Y2[NP * YP];
  const size_t nn = (size_t)n * n;
  const T* Eb = E + (size_t)b * nE * nn;
  T* xb = x + (size_t)b * N * n * m;
  const bool right = i + s < N, left = i >= s;
  load_y(Y0, xb + (size_t)i * n * m, n, m, c0, w, YP, NP);
  if (right) load_y(Y1, xb + (size_t)(i + s) * n * m, n, m, c0, w, YP, NP);
  if (left) load_y(Y2, xb + (size_t)(i - s) * n * m, n, m, c0, w, YP, NP);
  __syncthreads();
  // x_i -= R^T x_{i+s}, R = block (i+s, i)
  if (right) panel_sub_AtY<T, NB>(Y0, Eb + (size_t)(off + i / s) * nn, Y1, n, w, YP);
  __syncthreads();
  // x_i -= Lc x_{i-s}, Lc = block (i, i-s)
  if (left) panel_sub_AY<T, NB>(Y0, Eb + (size_t)(off + (i - s) / s) * nn, Y2, n, w, YP);
  __syncthreads();
  panel_upper_solve_T<T, NB>(Y0, L + ((size_t)b * N + i) * nn, n, w, YP);
  store_y(xb + (size_t)i * n * m, Y0, n, m, c0, w, YP);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T, int NB>
constexpr size_t potrf_blk_lds() {
  return sizeof(T) * ((size_t)(32 * NB) * (32 * NB + 1) + 64);
}
template <typename T, int NB>
constexpr size_t trsm_blk_lds() {
  return sizeof(T) * ((size_t)(NB * (NB + 1) / 2) * 32 * 33 + 32 * NB + (size_t)(32 * NB) * 33);
}

template <typename T, int NB>
static hipError_t set_blk_attrs() {
  hipError_t e;
  e = hipFuncSetAttribute((const void*)k_potrf_blk<T, NB, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)potrf_blk_lds<T, NB>());
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)k_potrf_blk<T, NB, 32 * NB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)potrf_blk_lds<T, NB>());
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)k_trsm_blk<T, NB, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)trsm_blk_lds<T, NB>());
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_trsm_blk<T, NB, 32 * NB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)trsm_blk_lds<T, NB>());
}

hipError_t init_blocked_kernels() {
  hipError_t e;
  if ((e = set_blk_attrs<float, 1>()) != hipSuccess) return e;
  if ((e = set_blk_attrs<float, 2>()) != hipSuccess) return e;
  if ((e = set_blk_attrs<float, 3>()) != hipSuccess) return e;
  if ((e = set_blk_attrs<float, 4>()) != hipSuccess) return e;
  if ((e = set_blk_attrs<double, 1>()) != hipSuccess) return e;
  if ((e = set_blk_attrs<double, 2>()) != hipSuccess) return e;
  if ((e = set_blk_attrs<double, 3>()) != hipSuccess) return e;
  if ((e = set_blk_attrs<double, 4>()) != hipSuccess) return e;
  return hipSuccess;
}

template <typename T, int NB, int NF>
static hipError_t factor_level_nb(T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, hipStream_t st) {
  const int nE = (int)off_diag_storage_len(N);
  const int Lv = recursive_iterations(N);
  const Level lv = level_info(N, lidx);
  if (lv.dim <= 0 || B <= 0) return hipSuccess;
  const size_t lds_potrf = potrf_blk_lds<T, NB>();
  const size_t lds_trsm = trsm_blk_lds<T, NB>();
  kt_begin(kKcPotrf, st);
  hipLaunchKernelGGL((k_potrf_blk<T, NB, NF>), dim3(lv.dim, B), dim3(kThreads), lds_potrf, st, L, mask, N, n, lv.s);
  kt_end(kKcPotrf, st);
  if (lidx == Lv - 1) return hipGetLastError();
  kt_begin(kKcTrsm, st);
  hipLaunchKernelGGL((k_trsm_blk<T, NB, NF>), dim3(lv.dim, 2 * NB, B), dim3(kThreads), lds_trsm, st, L, E, mask, N, n,
                     nE, lv.s, lv.off);
  kt_end(kKcTrsm, st);
  const int ntile = (n + kSTile - 1) / kSTile;
  const int ntri = ntile * (ntile + 1) / 2;
  const int nT = N / (2 * lv.s);
  int nF = 0;
  for (int t = 1; t < lv.dim; ++t)
    if (lv.s - 1 + 2 * lv.s * t + lv.s < N) ++nF;
  const int tasks = nT * ntri + nF * ntile * ntile;
  if (tasks > 0) {
    kt_begin(kKcSchur, st);
    hipLaunchKernelGGL((k_schur_blk<T, NF>), dim3(tasks, B), dim3(kThreads), 0, st, L, E, mask, N, n, nE, lv.s, lv.off,
                       lv.next_off, nT, ntile);
    kt_end(kKcSchur, st);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t factor_level_blocked(T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, hipStream_t st) {
  switch (n) {
    case 32: return factor_level_nb<T, 1, 32>(L, E, mask, B, N, n, lidx, st);
    case 64: return factor_level_nb<T, 2, 64>(L, E, mask, B, N, n, lidx, st);
    case 96: return factor_level_nb<T, 3, 96>(L, E, mask, B, N, n, lidx, st);
    case 128: return factor_level_nb<T, 4, 128>(L, E, mask, B, N, n, lidx, st);
    default: break;
  }
  switch ((n + 31) / 32) {
    case 1: return factor_level_nb<T, 1, 0>(L, E, mask, B, N, n, lidx, st);
    case 2: return factor_level_nb<T, 2, 0>(L, E, mask, B, N, n, lidx, st);
    case 3: return factor_level_nb<T, 3, 0>(L, E, mask, B, N, n, lidx, st);
    case 4: return factor_level_nb<T, 4, 0>(L, E, mask, B, N, n, lidx, st);
    default: return hipErrorInvalidValue;
  }
}

template <typename T, int NB, int NF>
static hipError_t forward_step_nb(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                  int step, hipStream_t st) {
  const int nE = (int)off_diag_storage_len(N);
  const int mch = (m + kRhsChunk - 1) / kRhsChunk;
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
  hipLaunchKernelGGL((k_fwd_blk<T, NB, NF>), dim3(items, mch, B), dim3(kThreads), 0, st, L, E, x, mask, N, n, m, nE, s, off,
                     step < 0 ? 1 : 0);
  kt_end(kKcForward, st);
  return hipGetLastError();
}

template <typename T>
hipError_t forward_step_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                int step, hipStream_t st) {
  switch (n) {
    case 32: return forward_step_nb<T, 1, 32>(L, E, x, mask, B, N, n, m, step, st);
    case 64: return forward_step_nb<T, 2, 64>(L, E, x, mask, B, N, n, m, step, st);
    case 96: return forward_step_nb<T, 3, 96>(L, E, x, mask, B, N, n, m, step, st);
    case 128: return forward_step_nb<T, 4, 128>(L, E, x, mask, B, N, n, m, step, st);
    default: break;
  }
  switch ((n + 31) / 32) {
    case 1: return forward_step_nb<T, 1, 0>(L, E, x, mask, B, N, n, m, step, st);
    case 2: return forward_step_nb<T, 2, 0>(L, E, x, mask, B, N, n, m, step, st);
    case 3: return forward_step_nb<T, 3, 0>(L, E, x, mask, B, N, n, m, step, st);
    case 4: return forward_step_nb<T, 4, 0>(L, E, x, mask, B, N, n, m, step, st);
    default: return hipErrorInvalidValue;
  }
}

template <typename T, int NB, int NF>
static hipError_t backward_level_nb(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                    int lidx, hipStream_t st) {
  const int nE = (int)off_diag_storage_len(N);
  const int mch = (m + kRhsChunk - 1) / kRhsChunk;
  const Level lv = level_info(N, lidx);
  if (lv.dim <= 0 || B <= 0 || mch <= 0) return hipSuccess;
  kt_begin(kKcBackward, st);
  hipLaunchKernelGGL((k_bwd_blk<T, NB, NF>), dim3(lv.dim, mch, B), dim3(kThreads), 0, st, L, E, x, mask, N, n, m, nE,
                     lv.s, lv.off);
  kt_end(kKcBackward, st);
  return hipGetLastError();
}

template <typename T>
hipError_t backward_level_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                  int lidx, hipStream_t st) {
  switch (n) {
    case 32: return backward_level_nb<T, 1, 32>(L, E, x, mask, B, N, n, m, lidx, st);
    case 64: return backward_level_nb<T, 2, 64>(L, E, x, mask, B, N, n, m, lidx, st);
    case 96: return backward_level_nb<T, 3, 96>(L, E, x, mask, B, N, n, m, lidx, st);
    case 128: return backward_level_nb<T, 4, 128>(L, E, x, mask, B, N, n, m, lidx, st);
    default: break;
  }
  switch ((n + 31) / 32) {
    case 1: return backward_level_nb<T, 1, 0>(L, E, x, mask, B, N, n, m, lidx, st);
    case 2: return backward_level_nb<T, 2, 0>(L, E, x, mask, B, N, n, m, lidx, st);
    case 3: return backward_level_nb<T, 3, 0>(L, E, x, mask, B, N, n, m, lidx, st);
    case 4: return backward_level_nb<T, 4, 0>(L, E, x, mask, B, N, n, m, lidx, st);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t factor_level_blocked<float>(float*, float*, const uint8_t*, int, int, int, int, hipStream_t);
template hipError_t factor_level_blocked<double>(double*, double*, const uint8_t*, int, int, int, int, hipStream_t);
template hipError_t forward_step_blocked<float>(const float*, const float*, float*, const uint8_t*, int, int, int, int,
                                                int, hipStream_t);
template hipError_t forward_step_blocked<double>(const double*, const double*, double*, const uint8_t*, int, int, int,
                                                 int, int, hipStream_t);
template hipError_t backward_level_blocked<float>(const float*, const float*, float*, const uint8_t*, int, int, int,
                                                  int, int, hipStream_t);
template hipError_t backward_level_blocked<double>(const double*, const double*, double*, const uint8_t*, int, int,
                                                   int, int, int, hipStream_t);

}  // namespace socu
