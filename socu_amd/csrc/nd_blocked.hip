// Blocked ND kernels for block size n <= 128 on gfx950: blocks are padded to
// NP = 32*NB (NB = ceil(n/32)) inside LDS/registers only; HBM accesses stay
// within the n x n blocks.  One or two kernel launches per phase per level; no
// atomics, so results are bitwise reproducible.
//
//   k_potrf_blk  one workgroup per node:  L_i <- chol(L_i), blocked
//                right-looking over 32x32 tiles (reference: potrf_blocked,
//                socu/blas/potrf.py:97-119).  Each diagonal tile is factored
//                and inverted by one wavefront (tile_chol_inv); panel and
//                trailing updates run on MFMA.  Then the full inverse Linv_i is
//                formed (MFMA) and stored transposed in the strict upper
//                triangle of L_i, which is not part of the factor (the
//                reference only ever reads tril(L)).
//   k_trsm_blk   one workgroup per (node, 32-wide chunk of R rows / Lc cols):
//                R <- R Linv^T, Lc <- Linv Lc as MFMA products with the stored
//                inverse (replaces trsm_rltn / trsm_llnn, trsm.py:94-120,
//                :202-228).
//   k_schur_blk  one workgroup per 64x64 output tile, MFMA, owner computes:
//                L_j -= R R^T + Lc^T Lc,  F -= R Lc (syrk.py, gemm.py).
//   k_sweep32, k_sweep_nb, k_apply_fwd  the sweeps, one launch per step and
//                direction: the block-times-panel update and the triangular
//                apply with the stored inverse in one kernel (k_apply_fwd:
//                the first forward level) (replaces K9-K16,
//                block_tridiag_solver.py:425-736).
#include <stdlib.h>

#include <hip/hip_ext.h>

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
// TS is the storage type of L; T the compute type (TS = float, T = double
// factors FP32 blocks with FP64 pivots and updates).
// One 64-lane wavefront for NB == 1 (the whole factor is wave 0's pivot
// chain: a workgroup of four would hold registers for three idle waves, and
// batches of many small blocks are throughput-bound); four otherwise.
constexpr int potrf_threads(int nb) { return nb == 1 ? 64 : kThreads; }

template <typename TS, int NB, int NF, typename T = TS>
__global__ __launch_bounds__(kThreads) void k_potrf_blk(TS* __restrict__ L, const uint8_t* __restrict__ mask, int N,
                                                       int n_, int s, const TS* __restrict__ Lsrc) {
  const int n = NF ? NF : n_;
  constexpr int NP = 32 * NB;
  constexpr int P = NP + 1;
  constexpr int TH = potrf_threads(NB);
  constexpr int WS = TH / 64 > 1 ? TH / 64 - 1 : 1;  // stride over waves 1.. (none when TH == 64)
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  TS* A = L + ((size_t)b * N + i) * (size_t)n * n;
  // out-of-place factor: the block is read from Lsrc (same layout), the
  // factor written to L
  const TS* Ain = (Lsrc != nullptr) ? Lsrc + ((size_t)b * N + i) * (size_t)n * n : A;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);
  T* dinv = S + NP * P;  // 1/L[r][r] for every row
  T* col = dinv + NP;
  T* ut = col + 128;  // 32 x 33: the last factored tile's inverse, transposed in full
  const int tid = threadIdx.x, wave = wave_id(), l = lane_id();

  const int nn = n * n;
  if constexpr (NF == NP) {
    // Unpadded block: only the lower tiles are read (every strictly upper
    // tile is written below before it is read).  Wave 0 reads diagonal tile
    // (0, 0) alone and starts its pivot chain at once; waves 1..3, idle in
    // that step, read the other lower tiles meanwhile (the step's barrier
    // publishes them).  Lane l reads row 2q + l/32, column l%32 of a tile.
    const int lr = l >> 5, lc = l & 31;
    if (wave == 0) {
      TS v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = Ain[(2 * q + lr) * NF + lc];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int r = 2 * q + lr;
        S[r * P + lc] = (lc <= r) ? (T)v[q] : T(0);
      }
    } else {
      constexpr int NT = NB * (NB + 1) / 2 - 1;  // lower tiles other than (0, 0)
      constexpr int PW = NT > 0 ? (NT + 2) / 3 : 1;
      TS v[PW][16];
#pragma unroll
      for (int k = 0; k < PW; ++k) {
        int ti = 1, tj = (wave - 1) + 3 * k;
        while (tj > ti) { tj -= ti + 1; ++ti; }
        if (ti < NB) {
#pragma unroll
          for (int q = 0; q < 16; ++q) v[k][q] = Ain[(32 * ti + 2 * q + lr) * NF + 32 * tj + lc];
        }
      }
#pragma unroll
      for (int k = 0; k < PW; ++k) {
        int ti = 1, tj = (wave - 1) + 3 * k;
        while (tj > ti) { tj -= ti + 1; ++ti; }
        if (ti < NB) {
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const int r = 2 * q + lr;
            S[(32 * ti + r) * P + 32 * tj + lc] = (ti != tj || lc <= r) ? (T)v[k][q] : T(0);
          }
        }
      }
    }
  } else {
    // The block is read as one contiguous n*n run (coalesced, unconditional
    // loads from clamped addresses) and scattered into the padded LDS square.
    constexpr int PER = NP * NP / TH;
    T v[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = tid + e * TH;
      v[e] = (T)Ain[min(idx, nn - 1)];
    }
    for (int idx = tid; idx < NP * NP; idx += TH) {
      const int r = idx / NP, c = idx - r * NP;
      S[r * P + c] = (r == c) ? T(1) : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = tid + e * TH;
      const int r = idx / n, c = idx - r * n;
      if (idx < nn) S[r * P + c] = (c <= r) ? v[e] : T(0);
    }
    __syncthreads();
  }

  // Off-diagonal tiles of the inverse:
  //   Linv_ij = -Dinv_i S_ij,  S_ij = sum_{k=j}^{i-1} L_ik Linv_kj  (j < i),
  //   Linv_jj = Dinv_j.
  // Linv_ij is kept transposed in the upper tile (j, i):
  // S[32j+c][32i+r] = Linv_ij[r][c]; S_ij is parked there first and then
  // overwritten in place.  Column j belongs to wavefront 1 + j, which forms
  // its blocks while wavefront 0 runs the pivot chain of the next diagonal
  // tile: in the step of tile kb it finishes Linv_{kb-1,j} and accumulates
  // S_{kb,j}, whose operands (L_{kb,k} for k < kb, Linv_kj for k < kb) are
  // final by then.
  auto inv_partial = [&](int i, int j) {  // S_ij -> upper tile (j, i)
    typename Mfma<T>::acc_t acc[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
    for (int k = j; k < i; ++k) {
      const T* Lik = S + 32 * i * P + 32 * k;
      if (k == j) {
        const T* Dj = S + 32 * j * P + 32 * j;
        wave_gemm_f<T>(
            acc, [&](int r, int kk) { return Lik[r * P + kk]; },
            [&](int kk, int c) { return (kk > c) ? Dj[c * P + kk] : ((kk == c) ? dinv[32 * j + c] : T(0)); }, 32);
      } else {
        const T* Ukj = S + 32 * j * P + 32 * k;  // holds Linv_kj transposed
        wave_gemm<T>(acc, Lik, P, 1, Ukj, 1, P, 32);
      }
    }
    T* Uji = S + 32 * j * P + 32 * i;
    acc_foreach<T>(acc, [&](int r, int c, T v) { Uji[c * P + r] = v; });
  };
  auto inv_finish = [&](int i, int j) {  // Linv_ij = -Dinv_i S_ij, in place
    T* Uji = S + 32 * j * P + 32 * i;
    typename Mfma<T>::acc_t inv[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
    const T* Di = S + 32 * i * P + 32 * i;
    wave_gemm_f<T>(
        inv, [&](int r, int kk) { return (r > kk) ? Di[kk * P + r] : ((r == kk) ? dinv[32 * i + r] : T(0)); },
        [&](int kk, int c) { return Uji[c * P + kk]; }, 32);
    acc_foreach<T>(inv, [&](int r, int c, T v) { Uji[c * P + r] = -v; });
  };

  // 32x32 tile (ti, tj) of S written back to the same tile of A (unpadded
  // blocks only): lane l stores row 2q + l/32, column l%32.
  auto store_tile = [&](int ti, int tj) {
    const int lr = l >> 5, lc = l & 31;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int r = 32 * ti + 2 * q + lr, c = 32 * tj + lc;
      A[r * n + c] = (TS)S[r * P + c];
    }
  };

  // tile (gi, gj) -= L_{gi,kb} L_{gj,kb}^T
  auto trail = [&](int kb, int gi, int gj) {
    typename Mfma<T>::acc_t acc[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
    wave_gemm<T>(acc, S + 32 * gi * P + 32 * kb, P, 1, S + 32 * gj * P + 32 * kb, 1, P, 32);
    T* C = S + 32 * gi * P + 32 * gj;
    acc_foreach<T>(acc, [&](int r, int c, T v) { C[r * P + c] -= v; });
  };

  // Two phases per diagonal tile kb, with look-ahead: wave 0 applies the
  // previous step's update to tile (kb, kb) only, then factors it; meanwhile
  // waves 1..3 apply the rest of that update (tiles (gi, gj), kb <= gj <= gi,
  // except (kb, kb)) and the inverse-block work.  Then every wave forms the
  // panel tiles (ti, kb) <- tile (ti, kb) Linv_kb^T.
  for (int kb = 0; kb < NB; ++kb) {
    T* D = S + 32 * kb * P + 32 * kb;
    if (wave == 0) {
      if (kb > 0) trail(kb - 1, kb, kb);
      tile_chol_inv<T>(D, P, dinv + 32 * kb, col, ut);
    } else {
      for (int j = wave - 1; j < NB - 1; j += WS) {
        if (j < kb - 1) inv_finish(kb - 1, j);
        if (j <= kb - 1) inv_partial(kb, j);
      }
      if (kb > 0) {
        const int m = NB - kb;
        const int cnt = m * (m + 1) / 2;  // triangle index 0 is (kb, kb)
        for (int t = wave; t < cnt; t += WS) {
          int ti = 0, q = t;
          while (q > ti) { q -= ti + 1; ++ti; }
          trail(kb - 1, kb + ti, kb + q);
        }
      }
      if constexpr (NF == NP) {
        // Write-back of the tiles final since the end of step kb - 1 (the
        // factor's tile column kb - 1 with the inverse of diagonal tile
        // kb - 1, and the inverse blocks Linv_{kb-2, j}^T), while wave 0 is in
        // its pivot chain.
        if (kb > 0) {
          const int nlow = NB - (kb - 1), nup = kb >= 2 ? kb - 2 : 0;
          for (int t = wave - 1; t < nlow + nup; t += WS) {
            if (t < nlow) store_tile(kb - 1 + t, kb - 1);
            else store_tile(t - nlow, kb - 2);
          }
        }
      }
    }
    __syncthreads();
    // panel: tile (ti, kb) <- tile (ti, kb) Linv^T
    for (int ti = kb + 1 + wave; ti < NB; ti += TH / 64) {
      T* X = S + 32 * ti * P + 32 * kb;
      typename Mfma<T>::acc_t acc[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
      const int li = l & 15, kq = l >> 4;
#pragma unroll
      for (int k = 0; k < 32; k += 4) {
        const int kk = k + kq;
        const T a0 = X[li * P + kk], a1 = X[(li + 16) * P + kk];
        // B(k, c) = Linv[c][k] = ut[k][c]
        const int c0 = li, c1 = li + 16;
        const T b0 = ut[kk * 33 + c0];
        const T b1 = ut[kk * 33 + c1];
        acc[0][0] = Mfma<T>::mma(a0, b0, acc[0][0]);
        acc[0][1] = Mfma<T>::mma(a0, b1, acc[0][1]);
        acc[1][0] = Mfma<T>::mma(a1, b0, acc[1][0]);
        acc[1][1] = Mfma<T>::mma(a1, b1, acc[1][1]);
      }
      acc_foreach<T>(acc, [&](int r, int c, T v) { X[r * P + c] = v; });
    }
    __syncthreads();
  }

  // The last row of inverse blocks (the others were formed during the
  // pivot chains above).
  if (wave > 0)
    for (int j = wave - 1; j < NB - 1; j += WS) inv_finish(NB - 1, j);
  __syncthreads();

  // Write-back: tril(L) is the factor; the strict upper triangle, which is
  // not part of the factor, carries the transposed inverse of the diagonal
  // block (U[k][r] = Linv[r][k], k < r) for the trsm and solve kernels.
  if constexpr (NF == NP) {
    // what the pivot steps have not written yet: diagonal tile NB-1 and the
    // inverse blocks Linv_{NB-2, j}^T, Linv_{NB-1, j}^T
    const int n2 = NB >= 2 ? NB - 2 : 0, cnt = 1 + n2 + (NB - 1);
    for (int t = wave; t < cnt; t += TH / 64) {
      if (t == 0) store_tile(NB - 1, NB - 1);
      else if (t - 1 < n2) store_tile(t - 1, NB - 2);
      else store_tile(t - 1 - n2, NB - 1);
    }
  } else {
    // one contiguous run
    for (int idx = tid; idx < nn; idx += TH) {
      const int r = idx / n, c = idx - r * n;
      A[idx] = (TS)S[r * P + c];
    }
  }
}

// ---------------------------------------------------------------------------
// One whole level-l node step for n == 32, by one wavefront per node:
//   L_i <- chol(L_i) with the transposed inverse in its strict upper triangle
//          (as k_potrf_blk),
//   R_i <- R_i Linv^T,  Lc_i <- Linv Lc_i          (as k_trsm_node),
//   F   <- -R_i Lc_i    the fill-in block (i+s, i-s) of the next level (the
//                       fill-in tasks of k_schur_small / k_schur_blk).
// The reference runs these as separate potrf, trsm and gemm passes over the
// level (socu/blas/potrf.py:97-119, socu/blas/trsm.py:94-120 and :202-228,
// block_tridiag_solver.py:797-813).  At n = 32 the pivot chain of the
// diagonal block is latency-bound and uses little of a compute unit, so the
// coupling blocks' loads, MFMA products and stores of one node run while the
// other wavefronts on the same SIMD are in their chains; R_i and Lc_i are
// read once (the separate kernels read them for the trsm and again for the
// fill-in) and the factor is not re-read for the trsm.
// R_i = block (i+s, i) exists when i + s < N, Lc_i = block (i, i-s) when
// i >= s, F when both do.  Layouts and storage as the separate kernels;
// Lsrc / Esrc (optional) are the out-of-place sources of L_i, R_i and Lc_i.
// With x (n x m blocks, optional) the node also takes its forward-sweep step
// y_i <- Linv y_i (k_apply_fwd); the updates y_j -= ... that level l sends to
// its targets are then k_schur_small's, so that a factor_and_solve needs no
// separate forward-sweep kernels.
// ---------------------------------------------------------------------------
constexpr size_t node32_lds(size_t esize) { return esize * (2 * 32 * 33 + 32 + 128); }

// (two wavefronts per SIMD: the register budget is held to 256, which the
// compiler meets without spills by keeping the MFMA accumulators in VGPRs)
template <typename T>
__global__ __launch_bounds__(64, 2) void k_node32(T* __restrict__ L, T* __restrict__ E, const uint8_t* __restrict__ mask,
                                              int N, int nE, int s, int off, int next_off,
                                              const T* __restrict__ Lsrc, const T* __restrict__ Esrc,
                                              T* __restrict__ x, int m) {
  constexpr int n = 32, P = 33;
  constexpr size_t nn = 32 * 32;
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  const bool hasR = i + s < N, hasC = i >= s;
  T* A = L + ((size_t)b * N + i) * nn;
  const T* Ain = (Lsrc != nullptr) ? Lsrc + (A - L) : A;
  T* Eb = E + (size_t)b * nE * nn;
  T* XR = hasR ? Eb + (size_t)(off + i / s) * nn : nullptr;        // block (i+s, i)
  T* XC = hasC ? Eb + (size_t)(off + (i - s) / s) * nn : nullptr;  // block (i, i-s)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);  // 32 x 33: the block, then R_i, Lc_i, R_i'
  T* ut = S + 32 * P;                       // 32 x 33: Linv transposed, then Lc_i'
  T* dinv = ut + 32 * P;
  T* col = dinv + 32;
  const int l = lane_id();
  const int lr = l >> 5, lc = l & 31;  // lane l moves row 2q + lr, column lc of a block

  {
    T v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = Ain[(2 * q + lr) * n + lc];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int r = 2 * q + lr;
      S[r * P + lc] = (lc <= r) ? v[q] : T(0);
    }
  }
  // R_i and Lc_i are loaded late in the pivot chain (pair step 20), when the
  // chain's column registers of the early steps are free, and land while the
  // chain finishes (earlier, the kernel needs more than the 256 registers
  // that let two wavefronts share a SIMD).
  T rv[16], cv[16];
  const T* XRin = hasR ? ((Esrc != nullptr) ? Esrc + (XR - E) : XR) : nullptr;
  const T* XCin = hasC ? ((Esrc != nullptr) ? Esrc + (XC - E) : XC) : nullptr;
  // y_i (forward sweep): up to 4 columns (128 entries, two per lane) are
  // loaded with R_i and Lc_i; wider ones at their step
  T* y = (x != nullptr) ? x + ((size_t)b * N + i) * (size_t)n * m : nullptr;
  const bool ypre = y != nullptr && m <= 4;
  T yv0 = T(0), yv1 = T(0);
  auto hook = [&](int k) {
    if (k == 20) {
      if (ypre) {
        yv0 = y[min(l, 32 * m - 1)];
        yv1 = y[min(l + 64, 32 * m - 1)];
      }
      if (hasR) {
#pragma unroll
        for (int q = 0; q < 16; ++q) rv[q] = XRin[(2 * q + lr) * n + lc];
      }
      if (hasC) {
#pragma unroll
        for (int q = 0; q < 16; ++q) cv[q] = XCin[(2 * q + lr) * n + lc];
      }
    }
  };
  tile_chol_inv<T>(S, P, dinv, col, ut, hook);
  // the factor (lower) with the transposed inverse (strict upper)
#pragma unroll
  for (int q = 0; q < 16; ++q) A[(2 * q + lr) * n + lc] = S[(2 * q + lr) * P + lc];
  if (y != nullptr) {
    // y_i <- Linv y_i, up to 32 columns at a time staged in S:
    // out[r][c] = sum_k ut[k][r] y[k][c]; lane l forms row l % 32 of the
    // columns of its half (l / 32)
    for (int c0 = 0; c0 < m; c0 += 32) {
      const int w = min(32, m - c0);
      if (ypre) {  // entries l and l + 64 of the 32 x m block
        if (l < 32 * m) S[(l / m) * P + l % m] = yv0;
        if (l + 64 < 32 * m) S[((l + 64) / m) * P + (l + 64) % m] = yv1;
      } else {
        for (int q = l; q < 32 * w; q += 64) {
          const int k = q / w, c = q - k * w;
          S[k * P + c] = y[(size_t)k * m + c0 + c];
        }
      }
      for (int c = lr; c < w; c += 2) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < 32; ++k) acc = fma(ut[k * P + lc], S[k * P + c], acc);
        y[(size_t)lc * m + c0 + c] = acc;
      }
    }
  }
  if (!hasR && !hasC) return;

  // (LDS operations of one wavefront complete in issue order: S and ut are
  // rewritten below only after the reads of their previous contents.)
  typename Mfma<T>::acc_t ar[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
  typename Mfma<T>::acc_t ac[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
  if (hasR) {
    // R' = R Linv^T: B(k, c) = Linv[c][k] = ut[k][c]
#pragma unroll
    for (int q = 0; q < 16; ++q) S[(2 * q + lr) * P + lc] = rv[q];
    wave_gemm<T>(ar, S, P, 1, ut, P, 1, 32);
    acc_foreach<T>(ar, [&](int r, int c, T v) { XR[r * n + c] = v; });
  }
  if (hasC) {
    // Lc' = Linv Lc: A(r, k) = Linv[r][k] = ut[k][r]
#pragma unroll
    for (int q = 0; q < 16; ++q) S[(2 * q + lr) * P + lc] = cv[q];
    wave_gemm<T>(ac, ut, 1, P, S, P, 1, 32);
    acc_foreach<T>(ac, [&](int r, int c, T v) { XC[r * n + c] = v; });
  }
  if (hasR && hasC) {
    // F = -R' Lc' from R' in S and Lc' in ut
    T* F = Eb + (size_t)(next_off + (i - s) / (2 * s)) * nn;
    acc_foreach<T>(ar, [&](int r, int c, T v) { S[r * P + c] = v; });
    acc_foreach<T>(ac, [&](int r, int c, T v) { ut[r * P + c] = v; });
    typename Mfma<T>::acc_t af[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
    wave_gemm<T>(af, S, P, 1, ut, P, 1, 32);
    acc_foreach<T>(af, [&](int r, int c, T v) { F[r * n + c] = -v; });
  }
}

// ---------------------------------------------------------------------------
// Inverses of the diagonal factor blocks from tril(L) alone, for solves whose
// factor may come from anywhere (the reference's sweeps read only tril(L),
// socu/blas/potrf.py:101-103).  One workgroup of NB wavefronts per block
// (grid (N, B)): W_i <- tril(L_i) with the transposed inverse in its strict
// upper triangle, W_i[k][r] = Linv_i[r][k] for k < r -- the layout k_potrf_blk
// leaves in L_i, which the sweep and apply kernels read.  L is only read.
//   wave w inverts diagonal tile w (tile_tri_inv);
//   band d = 1..NB-1: Linv_ij = -Dinv_i sum_{k=j}^{i-1} L_ik Linv_kj, i = j+d,
//   one wavefront per tile, MFMA.
// ---------------------------------------------------------------------------
template <typename T, int NB>
constexpr size_t trtri_blk_lds() {
  return sizeof(T) * ((size_t)(32 * NB) * (32 * NB + 1) + 32 * NB);
}

template <typename T, int NB, int NF>
__global__ __launch_bounds__(64 * NB) void k_trtri_blk(const T* __restrict__ L, T* __restrict__ W,
                                                      const uint8_t* __restrict__ mask, int N, int n_) {
  const int n = NF ? NF : n_;
  constexpr int NP = 32 * NB;
  constexpr int P = NP + 1;
  constexpr int TH = 64 * NB;
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  const int i = blockIdx.x;
  const size_t nn = (size_t)n * n;
  const T* A = L + ((size_t)b * N + i) * nn;
  T* Wb = W + ((size_t)b * N + i) * nn;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);
  T* dinv = S + NP * P;
  const int tid = threadIdx.x, wave = wave_id();

  // tril(L_i) into the padded square: zeros above the diagonal and outside
  // n, ones on the padded part of the diagonal; the block is read as one
  // contiguous run (upper entries are loaded and dropped).
  for (int idx = tid; idx < NP * NP; idx += TH) {
    const int r = idx / NP, c = idx - r * NP;
    T v = (r == c) ? T(1) : T(0);
    if (r < n && c < n) v = (c <= r) ? A[r * n + c] : T(0);
    S[r * P + c] = v;
  }
  __syncthreads();
  tile_tri_inv<T>(S + 32 * wave * P + 32 * wave, P, dinv + 32 * wave);
  __syncthreads();
  for (int d = 1; d < NB; ++d) {
    const int j = wave, ii = wave + d;
    if (ii < NB) {
      // S_ij -> upper tile (j, ii), transposed
      typename Mfma<T>::acc_t acc[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
      for (int k = j; k < ii; ++k) {
        const T* Lik = S + 32 * ii * P + 32 * k;
        const T* Ukj = S + 32 * j * P + 32 * k;  // Linv_kj transposed (k > j), or tile j's own inverse
        const T* dj = dinv + 32 * k;
        wave_gemm_f<T>(
            acc, [&](int r, int kk) { return Lik[r * P + kk]; },
            [&](int kk, int c) {
              if (k != j) return Ukj[c * P + kk];
              return (kk > c) ? Ukj[c * P + kk] : ((kk == c) ? dj[c] : T(0));
            },
            32);
      }
      T* Uji = S + 32 * j * P + 32 * ii;
      acc_foreach<T>(acc, [&](int r, int c, T v) { Uji[c * P + r] = v; });
      // Linv_ij = -Dinv_i S_ij, in place.  Only this wavefront touches tile
      // (j, ii), and the LDS operations of one wavefront complete in issue
      // order, so its stores above precede these reads and these reads the
      // stores below.
      const T* Di = S + 32 * ii * P + 32 * ii;
      const T* di = dinv + 32 * ii;
      typename Mfma<T>::acc_t inv[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
      wave_gemm_f<T>(
          inv, [&](int r, int kk) { return (r > kk) ? Di[kk * P + r] : ((r == kk) ? di[r] : T(0)); },
          [&](int kk, int c) { return Uji[c * P + kk]; }, 32);
      acc_foreach<T>(inv, [&](int r, int c, T v) { Uji[c * P + r] = -v; });
    }
    __syncthreads();  // band d is final before band d + 1 reads it
  }
  for (int idx = tid; idx < (int)nn; idx += TH) {
    const int r = idx / n, c = idx - r * n;
    Wb[idx] = S[r * P + c];
  }
}

// ---------------------------------------------------------------------------
// trsm with the explicit inverse left by k_potrf_blk in the strict upper
// triangle of L_i (U[k][r] = Linv[r][k] for k < r, 1/L[r][r] on the diagonal):
//   R task  (32 rows of R):   R  <- R Linv^T,   wave q: output columns 32q..
//   Lc task (32 cols of Lc):  Lc <- Linv Lc,    wave q: output rows 32q..
// Both are MFMA products with the upper tiles (kb, q), kb <= q, of U; no
// sequential substitution remains.
// ---------------------------------------------------------------------------
// Eight wavefronts: output tile q of the chunk is summed over kb <= q by two
// of them (kb below and from h(q) = (q + 1) / 2), which meet in LDS, so the
// longest product chain is half as long as with one wavefront per tile.
constexpr int kTrsmThreads = 512;

// (output tile q, half) of wavefront w: waves w and w + 4 share a SIMD, and
// these pairs carry at most three of the ten tile products of NB = 4
// (q of waves 0..7: 3 3 2 2 0 0 1 1; half: 0 1 1 0 1 0 0 1).
__device__ __forceinline__ void trsm_wave_task(int w, int& q, int& half) {
  q = (0x50AF >> (2 * w)) & 3;
  half = (0x96 >> w) & 1;
}
// k_trsm_blk splits only NB = 4 (n > 96); smaller blocks keep one wavefront
// per output tile (with NB <= 2 most of eight wavefronts would idle, and
// their workgroups, many per CU in batches, would hold twice the registers).
constexpr int trsm_blk_threads(int nb) { return nb == 4 ? kTrsmThreads : kThreads; }

template <typename T, int NB, int NF>
__global__ __launch_bounds__(kTrsmThreads) void k_trsm_blk(const T* __restrict__ L, T* __restrict__ E,
                                                      const uint8_t* __restrict__ mask, int N, int n_, int nE, int s,
                                                      int off, const T* __restrict__ Esrc) {
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
  const T* Xin = (Esrc != nullptr) ? Esrc + (X - E) : X;  // out-of-place: R / Lc read from Esrc
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Ut = reinterpret_cast<T*>(smem_raw);  // NT upper tiles (kb <= q) of U, 32 x TP each
  T* Xs = Ut + NT * 32 * TP;                // R task: 32 x (NP+1); Lc task: NP x TP
  constexpr int XP = NP + 1;
  const int tid = threadIdx.x, wave = wave_id();

  // Issue every load before storing any: upper tiles of U (diagonal entries
  // turned into reciprocals, lower parts of diagonal tiles zeroed) and the
  // 32-wide chunk of R or Lc.
  // (All loads are unconditional from clamped addresses and pinned only
  // after the last one is issued, so they are in flight together; the
  // masking and reciprocals follow.)
  constexpr int TT = trsm_blk_threads(NB);
  constexpr bool SPLIT = TT == kTrsmThreads;
  constexpr int PT = 32 * 32 / TT;
  T uv[NT][PT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    int q = 0, kb = t;
    while (kb > q) { kb -= q + 1; ++q; }  // t = q(q+1)/2 + kb, kb <= q
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = tid + e * TT;
      const int R = 32 * kb + (idx >> 5), C = 32 * q + (idx & 31);
      uv[t][e] = A[(size_t)min(R, n - 1) * n + min(C, n - 1)];
    }
  }
  constexpr int PZ = 32 * NP / TT;
  T zv[PZ];
#pragma unroll
  for (int e = 0; e < PZ; ++e) {
    const int idx = tid + e * TT;
    const int rr = idx / NP, kr = idx - rr * NP;  // R[c0+rr][kr], row reads
    const int kc = idx >> 5, cc = idx & 31;       // Lc[kc][c0+cc]
    const size_t ar = (size_t)(c0 + min(rr, w - 1)) * n + min(kr, n - 1);
    const size_t ac = (size_t)min(kc, n - 1) * n + c0 + min(cc, w - 1);
    zv[e] = Xin[isR ? ar : ac];
  }
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int e = 0; e < PT; ++e) pin_reg(uv[t][e]);
#pragma unroll
  for (int e = 0; e < PZ; ++e) pin_reg(zv[e]);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    int q = 0, kb = t;
    while (kb > q) { kb -= q + 1; ++q; }
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = tid + e * TT;
      const int R = 32 * kb + (idx >> 5), C = 32 * q + (idx & 31);
      const T g = uv[t][e];
      uv[t][e] = (R < n && C < n) ? ((R < C) ? g : ((R == C) ? T(1) / g : T(0))) : ((R == C) ? T(1) : T(0));
    }
  }
#pragma unroll
  for (int e = 0; e < PZ; ++e) {
    const int idx = tid + e * TT;
    const int rr = idx / NP, kr = idx - rr * NP;
    const int kc = idx >> 5, cc = idx & 31;
    const bool ok = isR ? (rr < w && kr < n) : (cc < w && kc < n);
    zv[e] = ok ? zv[e] : T(0);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = tid + e * TT;
      Ut[t * 32 * TP + (idx >> 5) * TP + (idx & 31)] = uv[t][e];
    }
#pragma unroll
  for (int e = 0; e < PZ; ++e) {
    const int idx = tid + e * TT;
    if (isR) Xs[(idx / NP) * XP + (idx % NP)] = zv[e];
    else Xs[(idx >> 5) * TP + (idx & 31)] = zv[e];
  }
  __syncthreads();

  int q = wave, half = 0;
  if constexpr (SPLIT) trsm_wave_task(wave, q, half);
  const int kb0 = half ? (q + 1) / 2 : 0, kb1 = (half || !SPLIT) ? q + 1 : (q + 1) / 2;
  typename Mfma<T>::acc_t acc[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
  if (q < NB) {
    for (int kb = kb0; kb < kb1; ++kb) {
      const T* Utile = Ut + (q * (q + 1) / 2 + kb) * 32 * TP;  // U[32kb.., 32q..]
      if (isR) {
        // out[r][c] += sum_k R[r][32kb+k] U[32kb+k][32q+c]
        wave_gemm<T>(acc, Xs + 32 * kb, XP, 1, Utile, TP, 1, 32);
      } else {
        // out[r][c] += sum_k Linv[32q+r][32kb+k] Lc[32kb+k][c],  Linv[32q+r][32kb+k] = U[32kb+k][32q+r]
        wave_gemm<T>(acc, Utile, 1, TP, Xs + 32 * kb * TP, TP, 1, 32);
      }
    }
  }
  // the upper half's sums go through the (now free) tiles of U
  __syncthreads();
  T* Red = Ut + q * 32 * TP;
  if (q < NB && half) acc_foreach<T>(acc, [&](int r, int c, T v) { Red[r * TP + c] = v; });
  __syncthreads();
  if (q < NB && !half) {
    if (isR) {
      acc_foreach<T>(acc, [&](int r, int c, T v) {
        const int C = 32 * q + c;
        if (r < w && C < n) X[(size_t)(c0 + r) * n + C] = SPLIT ? v + Red[r * TP + c] : v;
      });
    } else {
      acc_foreach<T>(acc, [&](int r, int c, T v) {
        const int R = 32 * q + r;
        if (c < w && R < n) X[(size_t)R * n + c0 + c] = SPLIT ? v + Red[r * TP + c] : v;
      });
    }
  }
}

// The same products with a few chunks per workgroup, for wide levels:
// k_trsm_blk reloads the node's upper tiles of U in each of its 2*NB
// workgroups (8x the block at n=128), which dominates its HBM traffic when
// every CU has work.  Here a workgroup stages U once and walks `cpb`
// consecutive chunks of the node's list (its R row chunks, then its Lc column
// chunks), the next chunk's loads in flight while the current one is
// multiplied.
template <typename T, int NB, int NF>
__global__ __launch_bounds__(kTrsmThreads) void k_trsm_node(const T* __restrict__ L, T* __restrict__ E,
                                                       const uint8_t* __restrict__ mask, int N, int n_, int nE, int s,
                                                       int off, int cpb, const T* __restrict__ Esrc) {
  const int n = NF ? NF : n_;
  constexpr int NP = 32 * NB;
  constexpr int TP = 33;
  constexpr int NT = NB * (NB + 1) / 2;
  constexpr int XP = NP + 1;
  constexpr int PT = 32 * 32 / kTrsmThreads;
  constexpr int PZ = 32 * NP / kTrsmThreads;
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int i = s - 1 + 2 * s * blockIdx.x;
  const bool hasR = i + s < N, hasC = i >= s;
  const int nbn = (n + 31) / 32;  // chunks per block
  const int all = (hasR ? nbn : 0) + (hasC ? nbn : 0);
  // this workgroup's chunks: [q0, q0 + total) of the node's list
  const int q0 = blockIdx.y * cpb;
  const int total = min(cpb, all - q0);
  if (total <= 0) return;
  const size_t nn = (size_t)n * n;
  const T* A = L + ((size_t)b * N + i) * nn;
  T* XR = hasR ? E + ((size_t)b * nE + off + i / s) * nn : nullptr;        // R_i: block (i+s, i)
  T* XC = hasC ? E + ((size_t)b * nE + off + (i - s) / s) * nn : nullptr;  // Lc_i: block (i, i-s)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Ut = reinterpret_cast<T*>(smem_raw);
  T* Xs = Ut + NT * 32 * TP;
  T* Red = Xs + (32 * XP > NP * TP ? 32 * XP : NP * TP);  // 4 tiles: the upper halves' sums
  const int tid = threadIdx.x, wave = wave_id();
  // chunk q: R rows 32c.. for q < nbn (if R exists), then Lc columns 32c..
  auto chunk = [&](int q, bool& isR, int& c0, int& w, T*& X) {
    isR = hasR && q < nbn;
    const int c = isR ? q : q - (hasR ? nbn : 0);
    c0 = 32 * c;
    w = min(32, n - c0);
    X = isR ? XR : XC;
  };
  // (out-of-place: R and Lc are read from Esrc, same layout)
  auto zload = [&](T* zv, bool isR, int c0, int w, const T* Xd) {
    const T* X = (Esrc != nullptr) ? Esrc + (Xd - E) : Xd;
#pragma unroll
    for (int e = 0; e < PZ; ++e) {
      const int idx = tid + e * kTrsmThreads;
      const int rr = idx / NP, kr = idx - rr * NP;  // R[c0+rr][kr], row reads
      const int kc = idx >> 5, cc = idx & 31;       // Lc[kc][c0+cc]
      const size_t ar = (size_t)(c0 + min(rr, w - 1)) * n + min(kr, n - 1);
      const size_t ac = (size_t)min(kc, n - 1) * n + c0 + min(cc, w - 1);
      zv[e] = X[isR ? ar : ac];
    }
  };
  auto zstore = [&](T* zv, bool isR, int w) {
#pragma unroll
    for (int e = 0; e < PZ; ++e) {
      const int idx = tid + e * kTrsmThreads;
      const int rr = idx / NP, kr = idx - rr * NP;
      const int kc = idx >> 5, cc = idx & 31;
      const bool ok = isR ? (rr < w && kr < n) : (cc < w && kc < n);
      const T v = ok ? zv[e] : T(0);
      if (isR) Xs[(idx / NP) * XP + (idx % NP)] = v;
      else Xs[(idx >> 5) * TP + (idx & 31)] = v;
    }
  };

  // U tiles and the first chunk: every load issued before any is used
  T uv[NT][PT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    int q = 0, kb = t;
    while (kb > q) { kb -= q + 1; ++q; }
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = tid + e * kTrsmThreads;
      const int R = 32 * kb + (idx >> 5), C = 32 * q + (idx & 31);
      uv[t][e] = A[(size_t)min(R, n - 1) * n + min(C, n - 1)];
    }
  }
  bool isR;
  int c0, w;
  T* X;
  chunk(q0, isR, c0, w, X);
  T zv[PZ];
  zload(zv, isR, c0, w, X);
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int e = 0; e < PT; ++e) pin_reg(uv[t][e]);
#pragma unroll
  for (int e = 0; e < PZ; ++e) pin_reg(zv[e]);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    int q = 0, kb = t;
    while (kb > q) { kb -= q + 1; ++q; }
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = tid + e * kTrsmThreads;
      const int R = 32 * kb + (idx >> 5), C = 32 * q + (idx & 31);
      const T g = uv[t][e];
      Ut[t * 32 * TP + (idx >> 5) * TP + (idx & 31)] =
          (R < n && C < n) ? ((R < C) ? g : ((R == C) ? T(1) / g : T(0))) : ((R == C) ? T(1) : T(0));
    }
  }
  zstore(zv, isR, w);
  __syncthreads();

  for (int q = 0; q < total; ++q) {
    const bool more = q + 1 < total;
    bool isRn = false;
    int c0n = 0, wn = 0;
    T* Xn = nullptr;
    if (more) {
      chunk(q0 + q + 1, isRn, c0n, wn, Xn);
      zload(zv, isRn, c0n, wn, Xn);  // in flight during the products below
    }
    {
      // output tile qq, summed over kb <= qq by two wavefronts (see
      // k_trsm_blk); the upper half's sums meet the lower half's in Red
      int qq, half;
      trsm_wave_task(wave, qq, half);
      const int kb0 = half ? (qq + 1) / 2 : 0, kb1 = half ? qq + 1 : (qq + 1) / 2;
      typename Mfma<T>::acc_t acc[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
      if (qq < NB)
        for (int kb = kb0; kb < kb1; ++kb) {
          const T* Utile = Ut + (qq * (qq + 1) / 2 + kb) * 32 * TP;  // U[32kb.., 32qq..]
          if (isR) wave_gemm<T>(acc, Xs + 32 * kb, XP, 1, Utile, TP, 1, 32);
          else wave_gemm<T>(acc, Utile, 1, TP, Xs + 32 * kb * TP, TP, 1, 32);
        }
      T* Rq = Red + qq * 32 * TP;
      if (qq < NB && half) acc_foreach<T>(acc, [&](int r, int c, T v) { Rq[r * TP + c] = v; });
      __syncthreads();
      if (qq < NB && !half) {
        if (isR) {
          acc_foreach<T>(acc, [&](int r, int c, T v) {
            const int C = 32 * qq + c;
            if (r < w && C < n) X[(size_t)(c0 + r) * n + C] = v + Rq[r * TP + c];
          });
        } else {
          acc_foreach<T>(acc, [&](int r, int c, T v) {
            const int R = 32 * qq + r;
            if (c < w && R < n) X[(size_t)R * n + c0 + c] = v + Rq[r * TP + c];
          });
        }
      }
    }
    if (more) {
#pragma unroll
      for (int e = 0; e < PZ; ++e) pin_reg(zv[e]);
      __syncthreads();  // every wave is done with the current chunk
      zstore(zv, isRn, wn);
      __syncthreads();
      isR = isRn;
      c0 = c0n;
      w = wn;
      X = Xn;
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

// Sequence of up to two products over K chunks: acc += sum_p A_p B_p (BF:
// on BF16 MFMA, float storage).
template <typename T, bool BF>
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
    if (active) schur_gemm<T, BF>(acc, As + 32 * wi, 1, kSP, Bs + 32 * wj, kSP, 1, kSK);
  }
}

// Workgroups are handed to the eight XCDs round-robin (XCD = workgroup index
// mod 8, gridDim.x a multiple of 8), and each XCD has its own L2.  Tasks are
// numbered group-major -- group g: the output tiles of target g, then those
// of fill-in g, which share E blocks 2g+1 (and read 2g, 2g+2) -- and XCD x
// takes the contiguous range [x * per, (x + 1) * per) of that order, so the
// tiles that re-read one operand block run on one XCD at about the same
// time.  Slots past the last task (and the missing fill-in of the last
// group) return at once.
__host__ __device__ constexpr int schur_blk_groups(int nT, int nF) { return nT > nF ? nT : nF; }
__host__ __device__ constexpr int schur_blk_per_xcd(int nT, int nF, int ntile) {
  return (schur_blk_groups(nT, nF) * (ntile * (ntile + 1) / 2 + ntile * ntile) + 7) / 8;
}

template <typename T, int NF, bool BF = false>
__global__ __launch_bounds__(kThreads) void k_schur_blk(T* __restrict__ L, T* __restrict__ E,
                                                       const uint8_t* __restrict__ mask, int N, int n_, int nE, int s,
                                                       int off, int next_off, int nT, int nF, int ntile,
                                                       const T* __restrict__ Lsrc) {
  const int n = NF ? NF : n_;
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  __shared__ T As[kSK * kSP];
  __shared__ T Bs[kSK * kSP];
  const int ntri = ntile * (ntile + 1) / 2;
  const int gs = ntri + ntile * ntile;
  int task;
  bool target;
  {
    const int per = schur_blk_per_xcd(nT, nF, ntile);
    const int lg = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    const int g = lg / gs, q = lg - g * gs;
    target = q < ntri;
    if (target ? g >= nT : g >= nF) return;
    task = target ? g * ntri + q : nT * ntri + g * ntile * ntile + (q - ntri);
  }
  const size_t nn = (size_t)n * n;
  T* Eb = E + (size_t)b * nE * nn;
  const int wave = wave_id();
  const int wi = wave >> 1, wj = wave & 1;
  typename Mfma<T>::acc_t acc[2][2] = {{acc_zero<T>(), acc_zero<T>()}, {acc_zero<T>(), acc_zero<T>()}};
  T* C;
  int tr, tc;
  bool active = true;
  if (target) {
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
    products64<T, BF>(acc, active, (j + s < N) ? 2 : 1, Am, Am, form, n, tr * kSTile, tc * kSTile, As, Bs);
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
    products64<T, BF>(acc, active, 1, Am, Bm, form, n, tr * kSTile, tc * kSTile, As, Bs);
  }
  if (!active) return;
  const int rb = tr * kSTile + 32 * wi, cb = tc * kSTile + 32 * wj;
  if (target) {
    // (out-of-place: the target is read from Lsrc, same layout)
    const T* Cin = (Lsrc != nullptr) ? Lsrc + (C - L) : C;
    acc_foreach<T>(acc, [&](int r, int c, T v) {
      const int rr = rb + r, cc = cb + c;
      if (rr < n && cc < n) C[(size_t)rr * n + cc] = Cin[(size_t)rr * n + cc] - v;
    });
  } else {
    // the fill-in block receives its only update here: F = -R Lc, whatever
    // it held on entry (the reference's F -= R Lc on a zeroed block)
    acc_foreach<T>(acc, [&](int r, int c, T v) {
      const int rr = rb + r, cc = cb + c;
      if (rr < n && cc < n) C[(size_t)rr * n + cc] = -v;
    });
  }
}

// ---------------------------------------------------------------------------
// Schur-complement updates for levels with few nodes.  k_schur_blk walks K in
// chunks, one global round trip per chunk, which at the upper levels (a
// handful of tiles) is the whole cost.  Here a task is a 32x32 output tile
// (4x more tasks) and issues every load it needs at once: the 32-wide panels
// of its operands (all of K) and its own C values, then one LDS staging and
// the MFMA products.  Panels are kept as P[i][k] (pitch NP + 1): row panels
// are 32 rows of a block, column panels 32 of its columns, transposed.
// ---------------------------------------------------------------------------
// four panels of 32 x (NP+1).  (The forward sweep of diagonal tiles stages
// the two neighbours' right-hand sides, NP x kSmallYc each, and 32 rows of the
// target's, 32 x kSmallYc, in panel 1, which diagonal tiles do not use.)
constexpr int kSmallYc = 8;
static_assert(2 * 32 * kSmallYc + 32 * kSmallYc <= 32 * 33, "right-hand sides fit in one panel");
template <typename T, int NB>
constexpr size_t schur_small_lds() {
  return sizeof(T) * 4 * 32 * (size_t)(32 * NB + 1);
}

template <typename T, int NB>
__device__ __forceinline__ void panel_load(T* v, const T* __restrict__ M, bool cols, int i0, int n) {
  constexpr int NP = 32 * NB;
  constexpr int PE = 32 * NP / kThreads;
#pragma unroll
  for (int e = 0; e < PE; ++e) {
    const int idx = threadIdx.x + e * kThreads;
    int i, k;
    if (!cols) {  // row panel: contiguous along k
      i = idx / NP;
      k = idx - i * NP;
    } else {      // column panel: 32 contiguous entries of each row k
      k = idx >> 5;
      i = idx & 31;
    }
    const int rr = cols ? k : i0 + i, cc = cols ? i0 + i : k;
    v[e] = M[min(rr, n - 1) * n + min(cc, n - 1)];
  }
}

template <typename T, int NB>
__device__ __forceinline__ void panel_store(T* P, const T* v, bool cols, int i0, int n) {
  constexpr int NP = 32 * NB;
  constexpr int PE = 32 * NP / kThreads;
#pragma unroll
  for (int e = 0; e < PE; ++e) {
    const int idx = threadIdx.x + e * kThreads;
    int i, k;
    if (!cols) {
      i = idx / NP;
      k = idx - i * NP;
    } else {
      k = idx >> 5;
      i = idx & 31;
    }
    P[i * (NP + 1) + k] = (i0 + i < n && k < n) ? v[e] : T(0);
  }
}

// With x (optional; n x m blocks of right-hand sides) the diagonal-tile
// tasks of a target j also apply level l's forward-sweep update to rows
// 32 tr .. of y_j (the forward sweep's update):
//   y_j -= R_{j-s} y_{j-s} + Lc_{j+s}^T y_{j+s},
// with the operand panels the Schur product has staged: rows 32 tr.. of R and
// columns 32 tr.. of Lc.  The nodes' y must have had their level-l step
// y <- Linv y (k_node32 does it).
template <typename T, int NB, int NF, bool BF = false>
__global__ __launch_bounds__(kThreads) void k_schur_small(T* __restrict__ L, T* __restrict__ E,
                                                         const uint8_t* __restrict__ mask, int N, int n_, int nE, int s,
                                                         int off, int next_off, int nT, const T* __restrict__ Lsrc,
                                                         T* __restrict__ x, int m) {
  const int n = NF ? NF : n_;
  constexpr int NP = 32 * NB, PP = NP + 1;
  constexpr int PE = 32 * NP / kThreads;
  constexpr int NTRI = NB * (NB + 1) / 2;
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Pn = reinterpret_cast<T*>(smem_raw);  // four panels of 32 x PP
  const size_t nn = (size_t)n * n;
  T* Eb = E + (size_t)b * nE * nn;
  int task = blockIdx.x;
  const T* M[4];
  bool cols[4];
  int i0[4];
  int np;
  T* C;
  int tr, tc;
  bool diag;
  int jt = -1;  // the target (target tasks only)
  if (task < nT * NTRI) {
    // L_j -= R R^T + Lc^T Lc  (R = block (j, j-s), Lc = block (j+s, j))
    const int u = task / NTRI;
    int q = task - u * NTRI;
    tr = 0;
    while (q > tr) { q -= tr + 1; ++tr; }
    tc = q;
    const int j = 2 * s * (u + 1) - 1;
    jt = j;
    C = L + ((size_t)b * N + j) * nn;
    const T* R = Eb + (size_t)(off + (j - s) / s) * nn;
    const bool hasC = j + s < N;
    const T* Lc = hasC ? Eb + (size_t)(off + j / s) * nn : R;
    M[0] = R;  M[1] = R;  M[2] = Lc;  M[3] = Lc;
    cols[0] = false; cols[1] = false; cols[2] = true; cols[3] = true;
    i0[0] = 32 * tr; i0[1] = 32 * tc; i0[2] = 32 * tr; i0[3] = 32 * tc;
    np = hasC ? 2 : 1;
    diag = tr == tc;
  } else {
    // fill-in F -= R Lc  (node i: R = block (i+s, i), Lc = block (i, i-s))
    task -= nT * NTRI;
    const int f = task / (NB * NB);
    const int q = task - f * NB * NB;
    tr = q / NB;
    tc = q - tr * NB;
    const int i = s - 1 + 2 * s * (f + 1);
    const T* R = Eb + (size_t)(off + i / s) * nn;
    const T* Lc = Eb + (size_t)(off + (i - s) / s) * nn;
    C = Eb + (size_t)(next_off + (i - s) / (2 * s)) * nn;
    M[0] = R;  M[1] = Lc;  M[2] = R;  M[3] = Lc;
    cols[0] = false; cols[1] = true; cols[2] = false; cols[3] = true;
    i0[0] = 32 * tr; i0[1] = 32 * tc; i0[2] = 32 * tr; i0[3] = 32 * tc;
    np = 1;
    diag = false;
  }
  const int l = lane_id(), wave = wave_id();
  const int wi = wave >> 1, wj = wave & 1;
  // forward sweep (diagonal tiles of targets, with x): rows 32 tr .. of y_j
  // -= R_{j-s} y_{j-s} + Lc_{j+s}^T y_{j+s}, kSmallYc columns at a time
  const bool ysw = x != nullptr && diag;
  const size_t nm = (size_t)n * m;
  const T* ya = nullptr;
  const T* yc = nullptr;
  T* yj = nullptr;
  if (ysw) {
    T* xb = x + (size_t)b * N * nm;
    ya = xb + (size_t)(jt - s) * nm;
    yc = (np == 2) ? xb + (size_t)(jt + s) * nm : ya;
    yj = xb + (size_t)jt * nm;
  }
  constexpr int YA = 32 * NB * kSmallYc / kThreads;  // neighbour entries per thread
  T yav[YA], ycv[YA], yjv = T(0);
  // (unconditional loads from clamped addresses; masked when staged)
  auto yload = [&](int c0, int w) {
#pragma unroll
    for (int e = 0; e < YA; ++e) {
      const int q = threadIdx.x + e * kThreads, k = q / w, c = q - k * w;
      yav[e] = ya[(size_t)min(k, n - 1) * m + c0 + c];
      ycv[e] = yc[(size_t)min(k, n - 1) * m + c0 + c];
    }
    const int q = threadIdx.x, i = q / w, c = q - i * w;
    yjv = yj[(size_t)min(32 * tr + i, n - 1) * m + c0 + c];
  };
  // every load of the task at once: the first columns of the right-hand
  // sides, the panels, then this wave's C values.  (Diagonal tiles: panels 1
  // and 3 equal panels 0 and 2 and are neither loaded nor stored.)
  if (ysw) yload(0, min(kSmallYc, m));
  T v[4][PE];
#pragma unroll
  for (int p = 0; p < 4; ++p)
    if (p < 2 * np && !(diag && (p & 1))) panel_load<T, NB>(v[p], M[p], cols[p], i0[p], n);
  // C values: the target's (the fill-in block receives its only update here,
  // F = -R Lc, so its entry contents are not read)
  T cv[4] = {T(0), T(0), T(0), T(0)};
  if (blockIdx.x < nT * NTRI) {
    // (out-of-place: the target is read from Lsrc, same layout)
    const T* Cin = (Lsrc != nullptr) ? Lsrc + (C - L) : C;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = 32 * tr + 16 * wi + Mfma<T>::row(l, r), cc = 32 * tc + 16 * wj + Mfma<T>::col(l, r);
      cv[r] = Cin[min(rr, n - 1) * n + min(cc, n - 1)];
    }
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int e = 0; e < PE; ++e) pin_reg(v[p][e]);
#pragma unroll
  for (int r = 0; r < 4; ++r) pin_reg(cv[r]);
#pragma unroll
  for (int p = 0; p < 4; ++p)
    if (p < 2 * np && !(diag && (p & 1))) panel_store<T, NB>(Pn + p * 32 * PP, v[p], cols[p], i0[p], n);
  const bool hasC = np == 2;
  T* Ya = Pn + 32 * PP;  // (panel 1) NP x kSmallYc
  T* Yc = Ya + NP * kSmallYc;
  T* Yj = Yc + NP * kSmallYc;  // 32 x kSmallYc
  auto ystage = [&](int w) {
#pragma unroll
    for (int e = 0; e < YA; ++e) {
      const int q = threadIdx.x + e * kThreads, k = q / w, c = q - k * w;
      if (q < NP * w) {
        Ya[k * kSmallYc + c] = (k < n) ? yav[e] : T(0);
        Yc[k * kSmallYc + c] = (hasC && k < n) ? ycv[e] : T(0);
      }
    }
    if ((int)threadIdx.x < 32 * w) Yj[(threadIdx.x / w) * kSmallYc + threadIdx.x % w] = yjv;
  };
  if (ysw) ystage(min(kSmallYc, m));  // published by the panels' barrier
  __syncthreads();
  typename Mfma<T>::acc_t acc = acc_zero<T>();
  for (int p = 0; p < np; ++p)
    schur_gemm16<T, BF>(acc, Pn + (2 * p) * 32 * PP + 16 * wi * PP, PP, 1,
                        Pn + (diag ? 2 * p : 2 * p + 1) * 32 * PP + 16 * wj * PP, 1, PP, NP);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int rr = 32 * tr + 16 * wi + Mfma<T>::row(l, r), cc = 32 * tc + 16 * wj + Mfma<T>::col(l, r);
    if (rr < n && cc < n && (!diag || cc <= rr)) C[rr * n + cc] = cv[r] - acc[r];
  }
  if (!ysw) return;
  const T* P0 = Pn;               // P0[i][k] = R[32 tr + i][k]
  const T* P2 = Pn + 2 * 32 * PP;  // P2[i][k] = Lc[k][32 tr + i]
  const int tid = threadIdx.x;
  for (int c0 = 0; c0 < m; c0 += kSmallYc) {
    const int w = min(kSmallYc, m - c0);
    if (c0 > 0) {
      __syncthreads();  // every wave is done with the previous columns
      yload(c0, w);
      ystage(w);
      __syncthreads();
    }
    // one output per thread, (i, c) = (tid / w, tid % w): consecutive
    // threads store consecutive entries of y_j's rows
    if (tid < 32 * w) {
      const int i = tid / w, c = tid - i * w;
      T a0 = T(0), a1 = T(0);
#pragma unroll 8
      for (int k = 0; k < NP; k += 2) {
        a0 = fma(P0[i * PP + k], Ya[k * kSmallYc + c], a0);
        a1 = fma(P0[i * PP + k + 1], Ya[(k + 1) * kSmallYc + c], a1);
      }
      if (hasC) {
#pragma unroll 8
        for (int k = 0; k < NP; k += 2) {
          a0 = fma(P2[i * PP + k], Yc[k * kSmallYc + c], a0);
          a1 = fma(P2[i * PP + k + 1], Yc[(k + 1) * kSmallYc + c], a1);
        }
      }
      const int r = 32 * tr + i;
      if (r < n) yj[(size_t)r * m + c0 + c] = Yj[i * kSmallYc + c] - (a0 + a1);
    }
  }
}

// ---------------------------------------------------------------------------
// solves, with the explicit inverses of the diagonal blocks (see k_potrf_blk)
//
// Forward step l (levels ascending), k_sweep32 / k_sweep_nb:
//   for every receiving block j = 2s(u+1)-1:  y_j -= R_{j-s} y_{j-s} + Lc_{j+s}^T y_{j+s},
//   then, for the nodes j of level l+1 among them:  y_j <- Linv_j y_j
//   (k_apply_fwd: the applies of level 0, which no update precedes)
// Backward level l (levels descending), k_sweep32 / k_sweep_nb:
//   for every node i:  x_i <- Linv_i^T (x_i - R_i^T x_{i+s} - Lc_i x_{i-s})
// ---------------------------------------------------------------------------
// Right-hand-side panels in LDS have pitch yp = min(m, kRhsChunk) + 1
// (dynamic shared memory sized by the launch: few right-hand sides leave room
// for more workgroups per compute unit).
__host__ __device__ constexpr int rhs_pitch(int m) { return (m < kRhsChunk ? m : kRhsChunk) + 1; }


// y_j <- Linv_j y_j for the nodes j = s-1+2s*t of one level.
// out[r] = y[r] / L[r][r] + sum_{k<r} U[k][r] y[k]; thread (r, part) reads
// U[k][r] for its K part (coalesced in r); parts are summed through LDS.
template <typename T, int NB, int NF>
__global__ __launch_bounds__(kThreads) void k_apply_fwd(const T* __restrict__ L, T* __restrict__ x,
                                                       const uint8_t* __restrict__ mask, int N, int n_, int m,
                                                       int s) {
  const int n = NF ? NF : n_;
  constexpr int NP = 32 * NB;
  constexpr int PARTS = kThreads / NP;
  constexpr int KP = (NP + PARTS - 1) / PARTS;
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int j = s - 1 + 2 * s * blockIdx.x;
  const int c0 = blockIdx.y * kRhsChunk;
  const int w = min(kRhsChunk, m - c0);
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  const T* U = L + ((size_t)b * N + j) * nn;
  T* y = x + ((size_t)b * N + j) * nm;
  const int yp = rhs_pitch(m);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Y = reinterpret_cast<T*>(smem_raw);  // NP x yp
  T* Pt = Y + NP * yp;                     // PARTS x NP x yp
  const int tid = threadIdx.x;
  const int r = tid % NP, part = tid / NP;
  const bool act = part < PARTS;
  // All loads are unconditional (clamped addresses) and issued before any is
  // consumed: pinning them to registers only after the last one is issued
  // keeps the compiler from turning "cond ? load : 0" into one branch (and
  // one memory round trip) per load.  Masking follows.
  T a[KP];
#pragma unroll
  for (int e = 0; e < KP; ++e) a[e] = U[(size_t)min(part * KP + e, n - 1) * n + min(r, n - 1)];
  T dg = U[(size_t)min(r, n - 1) * n + min(r, n - 1)];
#pragma unroll
  for (int e = 0; e < KP; ++e) pin_reg(a[e]);
  pin_reg(dg);
#pragma unroll
  for (int e = 0; e < KP; ++e) {
    const int k = part * KP + e;
    a[e] = (act && k < r && r < n) ? a[e] : T(0);
  }
  for (int q = tid; q < NP * w; q += kThreads) {
    const int k = q / w, c = q - k * w;
    Y[k * yp + c] = (k < n) ? y[(size_t)k * m + c0 + c] : T(0);
  }
  __syncthreads();
  if (act) {
    for (int c = 0; c < w; ++c) {
      T acc = T(0);
#pragma unroll
      for (int e = 0; e < KP; ++e) {
        const int k = part * KP + e;
        if (k < NP) acc += a[e] * Y[k * yp + c];
      }
      Pt[(part * NP + r) * yp + c] = acc;
    }
  }
  __syncthreads();
  if (part == 0 && r < n) {
    for (int c = 0; c < w; ++c) {
      T acc = Y[r * yp + c] / dg;
      for (int p = 0; p < PARTS; ++p) acc += Pt[(p * NP + r) * yp + c];
      y[(size_t)r * m + c0 + c] = acc;
    }
  }
}


// One sweep step for n == 32 in one launch, one workgroup per (block, 16
// right-hand sides), both products on 16x16 MFMA tiles (waves 0 and 1 take
// rows 0-15 and 16-31, the right-hand sides are the 16 columns).
// Backward (FWD = false), level l, node i (the reference's backward
// substitution, block_tridiag_solver.py:581-643; update and apply in one):
//   t   = x_i - R_i^T x_{i+s} - Lc_i x_{i-s}     (R_i = block (i+s, i) when
//         i + s < N, Lc_i = block (i, i-s) when i >= s),
//   x_i = Linv_i^T t = t / L_i[r][r] + U t       (U: the transposed inverse in
//         the strict upper triangle of L_i).
// Forward (FWD = true), step l, target j = 2s(u+1)-1 (the reference's forward
// substitution, block_tridiag_solver.py:425-485; the update of step l and,
// for the targets that are level-(l+1) nodes (u even), the apply of level
// l+1 in one):
//   t   = y_j - R_{j-s} y_{j-s} - Lc_{j+s}^T y_{j+s}   (R_{j-s} = block
//         (j, j-s), Lc_{j+s} = block (j+s, j) when j + s < N),
//   y_j = t, or Linv_j t = t / L_j[r][r] + U^T t for u even.
// All three blocks and the three right-hand-side panels are loaded at once
// and staged in LDS (32 x 33 tiles; panels of pitch rhs_pitch(m)).  Columns
// at or past w read whatever follows in LDS (a guard keeps those reads in
// the allocation): they only reach output columns that are not stored.
template <typename T>
constexpr size_t sweep32_lds(int m) {
  return sizeof(T) * (3 * 32 * 33 + 3 * 32 * (size_t)rhs_pitch(m) + 16 + 32);
}

template <typename T, bool FWD>
__global__ __launch_bounds__(kThreads) void k_sweep32(const T* __restrict__ L, const T* __restrict__ E,
                                                     T* __restrict__ x, const uint8_t* __restrict__ mask, int N, int m,
                                                     int nE, int s, int off, int apply) {
  constexpr int n = 32, P = 33;
  constexpr size_t nn = 32 * 32;
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int c0 = blockIdx.y * kRhsChunk;
  const int w = min(kRhsChunk, m - c0);
  const int yp = rhs_pitch(m);
  const size_t nm = (size_t)n * m;
  const T* Eb = E + (size_t)b * nE * nn;
  T* xb = x + (size_t)b * N * nm;
  // Blocks used transposed (A(i, k) = M[k][i]) with panel Xp, and as they
  // stand (A(i, k) = M[i][k]) with panel Xm.
  int i;
  bool hasT, hasS, inv;
  const T* MT;
  const T* MS;
  const T* xp;
  const T* xm;
  if (!FWD) {
    i = s - 1 + 2 * s * blockIdx.x;
    hasT = i + s < N;  // R_i
    hasS = i >= s;     // Lc_i
    inv = true;
  } else {
    const int u = blockIdx.x;
    i = 2 * s * (u + 1) - 1;
    hasT = i + s < N;  // Lc_{j+s}
    hasS = true;       // R_{j-s}
    inv = apply != 0 && (u & 1) == 0;
  }
  const T* U = L + ((size_t)b * N + i) * nn;
  T* xi = xb + (size_t)i * nm;
  if (!FWD) {
    MT = hasT ? Eb + (size_t)(off + i / s) * nn : U;  // (U: a valid dummy address)
    MS = hasS ? Eb + (size_t)(off + (i - s) / s) * nn : U;
  } else {
    MT = hasT ? Eb + (size_t)(off + i / s) * nn : U;
    MS = Eb + (size_t)(off + (i - s) / s) * nn;
  }
  xp = hasT ? xb + (size_t)(i + s) * nm : xi;
  xm = hasS ? xb + (size_t)(i - s) * nm : xi;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Us = reinterpret_cast<T*>(smem_raw);  // 32 x 33 each
  T* Ts = Us + 32 * P;
  T* Ss = Ts + 32 * P;
  T* Xp = Ss + 32 * P;  // 32 x yp each; Xp later holds t
  T* Xm = Xp + 32 * yp;
  T* Xi = Xm + 32 * yp;
  T* dg = Xi + 32 * yp + 16;  // 1 / L[r][r] (after a 16-entry guard)
  const int tid = threadIdx.x, l = lane_id(), wave = wave_id();
  // every load at once (unconditional, from valid addresses), then the LDS
  // stores with the absent operands zeroed
  constexpr int PT = 32 * 32 / kThreads;
  T uv[PT], tv[PT], sv[PT];
#pragma unroll
  for (int e = 0; e < PT; ++e) {
    const int idx = tid + e * kThreads;
    uv[e] = U[idx];
    tv[e] = MT[idx];
    sv[e] = MS[idx];
  }
  // right-hand-side entries q = tid + kThreads e (row q / w, column q % w)
  constexpr int PQ = 32 * kRhsChunk / kThreads;
  T xpv[PQ], xmv[PQ], xiv[PQ];
#pragma unroll
  for (int e = 0; e < PQ; ++e) {
    const int q = tid + e * kThreads, qr = q / w, qc = q - qr * w;
    const size_t xo = (size_t)min(qr, 31) * m + c0 + qc;
    xpv[e] = xp[xo];
    xmv[e] = xm[xo];
    xiv[e] = xi[xo];
  }
#pragma unroll
  for (int e = 0; e < PT; ++e) {
    pin_reg(uv[e]);
    pin_reg(tv[e]);
    pin_reg(sv[e]);
  }
#pragma unroll
  for (int e = 0; e < PT; ++e) {
    const int idx = tid + e * kThreads;
    const int r = idx >> 5, c = idx & 31;
    Us[r * P + c] = uv[e];
    Ts[r * P + c] = hasT ? tv[e] : T(0);
    Ss[r * P + c] = hasS ? sv[e] : T(0);
    if (r == c) dg[r] = T(1) / uv[e];
  }
#pragma unroll
  for (int e = 0; e < PQ; ++e) {
    const int q = tid + e * kThreads, qr = q / w, qc = q - qr * w;
    if (q < 32 * w) {
      Xp[qr * yp + qc] = hasT ? xpv[e] : T(0);
      Xm[qr * yp + qc] = hasS ? xmv[e] : T(0);
      Xi[qr * yp + qc] = xiv[e];
    }
  }
  __syncthreads();
  // t = x_i - MT^T xp - MS xm: wave rt (0, 1) forms rows 16 rt .. 16 rt + 15
  const int rt = wave;
  typename Mfma<T>::acc_t acc = acc_zero<T>();
  if (rt < 2) {
    wave_gemm16<T>(acc, Ts + 16 * rt, 1, P, Xp, yp, 1, 32);
    wave_gemm16<T>(acc, Ss + 16 * rt * P, P, 1, Xm, yp, 1, 32);
  }
  if (!inv) {
    if (rt < 2)
      acc16_foreach<T>(acc, [&](int r, int c, T v) {
        if (c < w) xi[(size_t)(16 * rt + r) * m + c0 + c] = Xi[(16 * rt + r) * yp + c] - v;
      });
    return;
  }
  __syncthreads();  // every read of Xp done
  if (rt < 2)
    acc16_foreach<T>(acc, [&](int r, int c, T v) {
      if (c < w) Xp[(16 * rt + r) * yp + c] = Xi[(16 * rt + r) * yp + c] - v;
    });
  __syncthreads();
  // backward: x_i = A t, A(r, k) = U[r][k] (k > r), 1 / L[r][r] (k = r);
  // forward: y_j = A t, A(r, k) = U[k][r] (k < r), 1 / L[r][r] (k = r)
  if (rt < 2) {
    acc = acc_zero<T>();
    const int ii = l & 15, kq = l >> 4, row = 16 * rt + ii;
#pragma unroll
    for (int k = 0; k < 32; k += 4) {
      const int kk = k + kq;
      const T u = FWD ? Us[kk * P + row] : Us[row * P + kk];
      const T a = (FWD ? kk < row : kk > row) ? u : ((kk == row) ? dg[row] : T(0));
      acc = Mfma<T>::mma(a, Xp[kk * yp + ii], acc);
    }
    acc16_foreach<T>(acc, [&](int r, int c, T v) {
      if (c < w) xi[(size_t)(16 * rt + r) * m + c0 + c] = v;
    });
  }
}

// The same sweep step for the other block sizes n <= 128 (NB = 1..4; for
// NB >= 2 three staged blocks would not fit LDS): the block operands are read
// from global memory
// in the MFMA operand layout (each element is used once: the right-hand
// sides are the 16 columns of every product), the three panels and the
// intermediate t are staged in LDS; the triangular apply skips the zero
// 16-column blocks of Linv.  Four wavefronts take the NP/16 row tiles in turn,
// except in the FP64 backward sweep (sweepnb_wide): one wavefront per tile,
// which issues every block load of its tile, the apply's included, before its
// first MFMA.  Measured against the four-wavefront loop: the FP64 backward
// sweeps gain (n=64 N=1024 m=16 factor_and_solve 0.612 -> 0.605 ms, n=128
// N=512 m=8 1.160 -> 1.151 ms); the forward sweep, which runs beside the
// factor, slowed the factor's kernels more than it gained, and the FP32
// backward sweeps of the mixed solver's correction solves were slower.
// Unconditional loads from clamped addresses, masked after.
template <typename T, bool FWD>
constexpr bool sweepnb_wide() {
  return !FWD && sizeof(T) == 8;
}

template <typename T, int NB, bool FWD>
constexpr int sweepnb_threads() {
  return sweepnb_wide<T, FWD>() ? 64 * 2 * NB : kThreads;
}

template <typename T, int NB>
constexpr size_t sweepnb_lds(int m) {
  return sizeof(T) * (4 * (size_t)(32 * NB) * rhs_pitch(m) + 16 + 32 * NB);
}

template <typename T, int NB, int NF, bool FWD>
__global__ __launch_bounds__((sweepnb_threads<T, NB, FWD>())) void k_sweep_nb(const T* __restrict__ L,
                                                                        const T* __restrict__ E, T* __restrict__ x,
                                                                        const uint8_t* __restrict__ mask, int N,
                                                                        int n_, int m, int nE, int s, int off,
                                                                        int apply) {
  const int n = NF ? NF : n_;
  constexpr int NP = 32 * NB, RT = NP / 16, KS = NP / 4, kT = sweepnb_threads<T, NB, FWD>();
  const int b = blockIdx.z;
  if (masked_out(mask, b)) return;
  const int c0 = blockIdx.y * kRhsChunk;
  const int w = min(kRhsChunk, m - c0);
  const int yp = rhs_pitch(m);
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  const T* Eb = E + (size_t)b * nE * nn;
  T* xb = x + (size_t)b * N * nm;
  int i;
  bool hasT, hasS, inv;
  if (!FWD) {
    i = s - 1 + 2 * s * blockIdx.x;
    hasT = i + s < N;  // R_i, used transposed
    hasS = i >= s;     // Lc_i
    inv = true;
  } else {
    const int u = blockIdx.x;
    i = 2 * s * (u + 1) - 1;
    hasT = i + s < N;  // Lc_{j+s}, used transposed
    hasS = true;       // R_{j-s}
    inv = apply != 0 && (u & 1) == 0;
  }
  const T* U = L + ((size_t)b * N + i) * nn;
  T* xi = xb + (size_t)i * nm;
  const T* MT = hasT ? Eb + (size_t)(off + i / s) * nn : U;  // (U: a valid dummy address)
  const T* MS = hasS ? Eb + (size_t)(off + (i - s) / s) * nn : U;
  const T* xp = hasT ? xb + (size_t)(i + s) * nm : xi;
  const T* xm = hasS ? xb + (size_t)(i - s) * nm : xi;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Xp = reinterpret_cast<T*>(smem_raw);  // NP x yp each
  T* Xm = Xp + NP * yp;
  T* Xi = Xm + NP * yp;
  T* Tt = Xi + NP * yp;
  T* dg = Tt + NP * yp + 16;  // 1 / L[r][r] (after a 16-entry guard)
  const int tid = threadIdx.x, l = lane_id(), wave = wave_id();
  const int li = l & 15, kq = l >> 4;
  auto stage = [&]() {
    for (int q = tid; q < NP * w; q += kT) {
      const int k = q / w, c = q - k * w;
      const size_t xo = (size_t)min(k, n - 1) * m + c0 + c;
      const T vp = xp[xo], vm = xm[xo], vi = xi[xo];
      Xp[k * yp + c] = (hasT && k < n) ? vp : T(0);
      Xm[k * yp + c] = (hasS && k < n) ? vm : T(0);
      Xi[k * yp + c] = (k < n) ? vi : T(0);
    }
    for (int r = tid; r < NP; r += kT) {
      const T d = U[(size_t)min(r, n - 1) * (n + 1)];
      dg[r] = (r < n) ? T(1) / d : T(1);
    }
  };
  const int jb = min(li, w - 1);
  if constexpr (sweepnb_wide<T, FWD>()) {
    // backward, one wavefront per 16-row tile t (32-bit element offsets from the
    // wavefront-uniform block addresses)
    const int t = wave;
    const int ii = 16 * t + li;
    const unsigned ic = min(ii, n - 1);
    T a1[KS], a2[KS], a3[KS];
#pragma unroll
    for (int q = 0; q < KS; ++q) {
      const unsigned kc = min(4 * q + kq, n - 1);
      a1[q] = MT[kc * (unsigned)n + ic];
      a2[q] = MS[ic * (unsigned)n + kc];
    }
    stage();
    __syncthreads();
    // t = x_i - R_i^T x_{i+s} - Lc_i x_{i-s}
    typename Mfma<T>::acc_t acc = acc_zero<T>();
#pragma unroll
    for (int q = 0; q < KS; ++q) {
      const int kk = 4 * q + kq;
      pin_reg(a1[q]);
      pin_reg(a2[q]);
      const bool ok = kk < n && ii < n;
      acc = Mfma<T>::mma((hasT && ok) ? a1[q] : T(0), Xp[kk * yp + jb], acc);
      acc = Mfma<T>::mma((hasS && ok) ? a2[q] : T(0), Xm[kk * yp + jb], acc);
    }
    acc16_foreach<T>(acc, [&](int r, int c, T v) {
      if (c < w) Tt[(16 * t + r) * yp + c] = Xi[(16 * t + r) * yp + c] - v;
    });
    // the apply's operands (k >= 16 t), loaded while the other wavefronts
    // finish t; steps below 4 t load the entries of step 4 t again (the same
    // addresses: no further traffic) and are not used
#pragma unroll
    for (int q = 0; q < KS; ++q) {
      const unsigned kc = min(4 * max(q, 4 * t) + kq, n - 1);
      a3[q] = U[ic * (unsigned)n + kc];
    }
    __syncthreads();
    // x_i = A t, A(r, k) = U[r][k] (k > r), 1 / L[r][r] (k = r)
    acc = acc_zero<T>();
#pragma unroll
    for (int q = 0; q < KS; ++q) {
      if (q >= 4 * t) {  // wavefront-uniform
        const int kk = 4 * q + kq;
        pin_reg(a3[q]);
        const T a = (kk < n && ii < n) ? ((kk > ii) ? a3[q] : ((kk == ii) ? dg[ii] : T(0))) : T(0);
        acc = Mfma<T>::mma(a, Tt[kk * yp + jb], acc);
      }
    }
    acc16_foreach<T>(acc, [&](int r, int c, T v) {
      const int row = 16 * t + r;
      if (row < n && c < w) xi[(size_t)row * m + c0 + c] = v;
    });
    return;
  }
  stage();
  __syncthreads();
  // t = x_i - MT^T xp - MS xm
  for (int t = wave; t < RT; t += kT / 64) {
    const int ii = 16 * t + li, ic = min(ii, n - 1);
    typename Mfma<T>::acc_t acc = acc_zero<T>();
#pragma unroll 4
    for (int k = 0; k < NP; k += 4) {
      const int kk = k + kq, kc = min(kk, n - 1);
      const T a1 = MT[(size_t)kc * n + ic];
      const T a2 = MS[(size_t)ic * n + kc];
      const bool ok = kk < n && ii < n;
      acc = Mfma<T>::mma((hasT && ok) ? a1 : T(0), Xp[kk * yp + jb], acc);
      acc = Mfma<T>::mma((hasS && ok) ? a2 : T(0), Xm[kk * yp + jb], acc);
    }
    acc16_foreach<T>(acc, [&](int r, int c, T v) {
      const int row = 16 * t + r;
      if (c < w) {
        const T tv = Xi[row * yp + c] - v;
        if (inv) Tt[row * yp + c] = tv;
        else if (row < n) xi[(size_t)row * m + c0 + c] = tv;
      }
    });
  }
  if (!inv) return;
  __syncthreads();
  // backward: x_i = A t, A(r, k) = U[r][k] (k > r), 1 / L[r][r] (k = r);
  // forward: y_j = A t, A(r, k) = U[k][r] (k < r), 1 / L[r][r] (k = r)
  for (int t = wave; t < RT; t += kT / 64) {
    const int ii = 16 * t + li, ic = min(ii, n - 1);
    const int kb = FWD ? 0 : 16 * t, ke = FWD ? 16 * (t + 1) : NP;
    typename Mfma<T>::acc_t acc = acc_zero<T>();
#pragma unroll 4
    for (int k = kb; k < ke; k += 4) {
      const int kk = k + kq, kc = min(kk, n - 1);
      const T u = FWD ? U[(size_t)kc * n + ic] : U[(size_t)ic * n + kc];
      const bool strict = FWD ? kk < ii : kk > ii;
      const T a = (kk < n && ii < n) ? (strict ? u : ((kk == ii) ? dg[ii] : T(0))) : T(0);
      acc = Mfma<T>::mma(a, Tt[kk * yp + jb], acc);
    }
    acc16_foreach<T>(acc, [&](int r, int c, T v) {
      const int row = 16 * t + r;
      if (row < n && c < w) xi[(size_t)row * m + c0 + c] = v;
    });
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// dynamic LDS of the sweep kernels for m right-hand sides
template <typename T, int NB>
constexpr size_t apply_fwd_lds(int m) {
  return sizeof(T) * (1 + kThreads / (32 * NB)) * (size_t)(32 * NB) * rhs_pitch(m);
}

template <typename T, int NB>
constexpr size_t potrf_blk_lds() {
  return sizeof(T) * ((size_t)(32 * NB) * (32 * NB + 1) + 32 * NB + 128 + 32 * 33);
}
template <typename T, int NB>
constexpr size_t trsm_blk_lds() {
  // upper tiles + max(R chunk 32 x (NP+1), Lc chunk NP x 33)
  return sizeof(T) * ((size_t)(NB * (NB + 1) / 2) * 32 * 33 +
                      ((size_t)32 * (32 * NB + 1) > (size_t)(32 * NB) * 33 ? (size_t)32 * (32 * NB + 1)
                                                                             : (size_t)(32 * NB) * 33));
}

// k_trsm_node: k_trsm_blk's LDS and four 32 x 33 tiles for the half sums
template <typename T, int NB>
constexpr size_t trsm_node_lds() {
  return trsm_blk_lds<T, NB>() + sizeof(T) * 4 * 32 * 33;
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
  e = hipFuncSetAttribute((const void*)k_trsm_blk<T, NB, 32 * NB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)trsm_blk_lds<T, NB>());
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)k_trsm_node<T, NB, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)trsm_node_lds<T, NB>());
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)k_trsm_node<T, NB, 32 * NB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)trsm_node_lds<T, NB>());
  const int lsw = (int)sweepnb_lds<T, NB>(kRhsChunk);  // the widest right-hand-side chunk
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_sweep_nb<T, NB, 0, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lsw);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_sweep_nb<T, NB, 0, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lsw);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_sweep_nb<T, NB, 32 * NB, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            lsw);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_sweep_nb<T, NB, 32 * NB, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            lsw);
  return e;
}

// Whether a level's trsm takes k_trsm_node, and with how many chunks per
// workgroup.  Measured on MI355X (N=512): for n > 96 and levels of at least
// 64 nodes, two chunks per workgroup cut the n=128 factor by 9 us (FP32) and
// 35 us (FP64); for n <= 64, or one workgroup per node, it was slower.
// SOCU_TRSM_NODE=<min nodes> and SOCU_TRSM_CPB=<chunks> override.
static int trsm_node_min(int nb) {
  static int v = -2;
  if (v == -2) {
    const char* e = getenv("SOCU_TRSM_NODE");
    v = e ? (atoi(e) < 1 ? 1 : atoi(e)) : -1;
  }
  if (v > 0) return v;
  return nb == 4 ? 64 : (1 << 30);
}

static int trsm_chunks_per_block() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("SOCU_TRSM_CPB");
    v = e ? atoi(e) : 2;
    if (v < 1) v = 1;
  }
  return v;
}

template <typename T, int NB>
static hipError_t set_schur_small_attrs() {
  const int lds = (int)schur_small_lds<T, NB>();
  hipError_t e = hipFuncSetAttribute((const void*)k_schur_small<T, NB, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_schur_small<T, NB, 32 * NB>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  if constexpr (sizeof(T) == 4) {  // BF16 products (float storage only)
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)k_schur_small<T, NB, 0, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lds);
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)k_schur_small<T, NB, 32 * NB, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  }
  return e;
}

// Whether a level's Schur updates (nodes = targets + fill-ins over the batch)
// take k_schur_small.  Its panel loads repeat each operand per output tile,
// which only costs where the level streams many blocks; measured on MI355X
// (N=512, 1024) the small kernel wins up to about 12 MiB of the level's
// blocks: n=128 FP64 factor 1207 -> 1093 us, FP32 875 -> 767 us, n=32 FP64
// 265 -> 197 us.  For n <= 32 it is taken at every level: k_schur_blk's
// 64x64 output tiles leave three of four quarters outside a 32x32 block
// (B=256 n=32 N=128 m=4 FP64 factor_and_solve: 2.36 -> 1.65 ms).
// SOCU_SCHUR_SMALL=<nodes> replaces the rule by a node count (0 disables the
// small kernel).
static bool use_schur_small(int nodes, int n, int esize) {
  static int env = -2;
  if (env == -2) {
    const char* e = getenv("SOCU_SCHUR_SMALL");
    env = e ? (atoi(e) < 0 ? 0 : atoi(e)) : -1;
  }
  if (env >= 0) return nodes <= env;
  if (n <= 32) return true;
  return (double)nodes * n * n * esize <= 12.0 * 1024 * 1024;
}

template <typename T, int NB>
static hipError_t set_trtri_attrs() {
  hipError_t e = hipFuncSetAttribute((const void*)k_trtri_blk<T, NB, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)trtri_blk_lds<T, NB>());
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_trtri_blk<T, NB, 32 * NB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)trtri_blk_lds<T, NB>());
}

template <int NB>
static hipError_t set_mixed_attrs() {
  hipError_t e = hipFuncSetAttribute((const void*)k_potrf_blk<float, NB, 0, double>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)potrf_blk_lds<double, NB>());
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_potrf_blk<float, NB, 32 * NB, double>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)potrf_blk_lds<double, NB>());
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
  if ((e = set_schur_small_attrs<float, 1>()) != hipSuccess) return e;
  if ((e = set_schur_small_attrs<float, 2>()) != hipSuccess) return e;
  if ((e = set_schur_small_attrs<float, 3>()) != hipSuccess) return e;
  if ((e = set_schur_small_attrs<float, 4>()) != hipSuccess) return e;
  if ((e = set_schur_small_attrs<double, 1>()) != hipSuccess) return e;
  if ((e = set_schur_small_attrs<double, 2>()) != hipSuccess) return e;
  if ((e = set_schur_small_attrs<double, 3>()) != hipSuccess) return e;
  if ((e = set_schur_small_attrs<double, 4>()) != hipSuccess) return e;
  if ((e = set_trtri_attrs<float, 1>()) != hipSuccess) return e;
  if ((e = set_trtri_attrs<float, 2>()) != hipSuccess) return e;
  if ((e = set_trtri_attrs<float, 3>()) != hipSuccess) return e;
  if ((e = set_trtri_attrs<float, 4>()) != hipSuccess) return e;
  if ((e = set_trtri_attrs<double, 1>()) != hipSuccess) return e;
  if ((e = set_trtri_attrs<double, 2>()) != hipSuccess) return e;
  if ((e = set_trtri_attrs<double, 3>()) != hipSuccess) return e;
  if ((e = set_trtri_attrs<double, 4>()) != hipSuccess) return e;
  if ((e = set_mixed_attrs<1>()) != hipSuccess) return e;
  if ((e = set_mixed_attrs<2>()) != hipSuccess) return e;
  if ((e = set_mixed_attrs<3>()) != hipSuccess) return e;
  if ((e = set_mixed_attrs<4>()) != hipSuccess) return e;
  return hipSuccess;
}

template <typename T, int NB, int NF>
static hipError_t factor_level_nb(T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, int prec,
                                  hipStream_t st, hipEvent_t after_potrf, hipEvent_t after_trsm,
                                  hipEvent_t before_trsm, const T* Lsrc, const T* Esrc, T* x = nullptr,
                                  int m = 0) {
  const int nE = (int)off_diag_storage_len(N);
  const int Lv = recursive_iterations(N);
  const Level lv = level_info(N, lidx);
  if (lv.dim <= 0 || B <= 0) return hipSuccess;
  const size_t lds_trsm = trsm_blk_lds<T, NB>();
  const bool bf16 = (prec & kPrecBf16Schur) != 0 && sizeof(T) == 4;
  // n == 32 in the storage precision: one fused node kernel (k_node32) per
  // level instead of potrf, trsm and the Schur kernel's fill-in tasks (not
  // when the trsm must wait for other work, nor with FP64 pivots for FP32
  // blocks or BF16 Schur products, which the separate kernels implement)
  // With x (fused_forward_applies), the level's forward-sweep steps run in
  // the same kernels: y_i <- Linv y_i in k_node32, the updates of the targets'
  // y in k_schur_small.
  const bool fused = NF == 32 && before_trsm == nullptr && !bf16 &&
                     !((prec & kPrecF64Pivots) != 0 && sizeof(T) == 4);
  T* const xf = fused ? x : nullptr;
  if (fused) {
    hipEvent_t ea = nullptr, eb = nullptr;
    if (kt_kernel_events(kKcPotrf, &ea, &eb))
      hipExtLaunchKernelGGL(k_node32<T>, dim3(lv.dim, B), dim3(64), node32_lds(sizeof(T)), st, ea, eb, 0, L, E, mask, N,
                            nE, lv.s, lv.off, lv.next_off, Lsrc, Esrc, xf, m);
    else
      hipLaunchKernelGGL(k_node32<T>, dim3(lv.dim, B), dim3(64), node32_lds(sizeof(T)), st, L, E, mask, N, nE, lv.s,
                         lv.off, lv.next_off, Lsrc, Esrc, xf, m);
    hipError_t e = hipSuccess;
    if (after_potrf != nullptr) e = hipEventRecord(after_potrf, st);
    if (e == hipSuccess && after_trsm != nullptr) e = hipEventRecord(after_trsm, st);
    if (e != hipSuccess) return e;
    if (lidx == Lv - 1) return hipGetLastError();
  } else {
    // under socu_profile_op the launch itself records its start and end
    // (events of hipExtLaunchKernel: the kernel's own execution interval)
    auto k = k_potrf_blk<T, NB, NF, T>;
    size_t lds_potrf = potrf_blk_lds<T, NB>();
    if ((prec & kPrecF64Pivots) != 0 && sizeof(T) == 4) {
      k = k_potrf_blk<T, NB, NF, double>;
      lds_potrf = potrf_blk_lds<double, NB>();
    }
    hipEvent_t ea = nullptr, eb = nullptr;
    if (kt_kernel_events(kKcPotrf, &ea, &eb))
      hipExtLaunchKernelGGL(k, dim3(lv.dim, B), dim3(potrf_threads(NB)), lds_potrf, st, ea, eb, 0, L, mask, N, n, lv.s,
                            Lsrc);
    else
      hipLaunchKernelGGL(k, dim3(lv.dim, B), dim3(potrf_threads(NB)), lds_potrf, st, L, mask, N, n, lv.s, Lsrc);
    if (after_potrf != nullptr) {
      const hipError_t e = hipEventRecord(after_potrf, st);
      if (e != hipSuccess) return e;
    }
    if (lidx == Lv - 1) return hipGetLastError();
    if (before_trsm != nullptr) {
      const hipError_t e = hipStreamWaitEvent(st, before_trsm, 0);
      if (e != hipSuccess) return e;
    }
    kt_begin(kKcTrsm, st);
    if (B * lv.dim >= trsm_node_min(NB)) {
      const int cpb = trsm_chunks_per_block();
      const int groups = (2 * NB + cpb - 1) / cpb;
      const size_t lds_node = trsm_node_lds<T, NB>();
      hipLaunchKernelGGL((k_trsm_node<T, NB, NF>), dim3(lv.dim, groups, B), dim3(kTrsmThreads), lds_node, st, L, E,
                         mask, N, n, nE, lv.s, lv.off, cpb, Esrc);
    }
    else
      hipLaunchKernelGGL((k_trsm_blk<T, NB, NF>), dim3(lv.dim, 2 * NB, B), dim3(trsm_blk_threads(NB)), lds_trsm, st, L,
                         E, mask, N, n, nE, lv.s, lv.off, Esrc);
    kt_end(kKcTrsm, st);
    if (after_trsm != nullptr) {
      const hipError_t e = hipEventRecord(after_trsm, st);
      if (e != hipSuccess) return e;
    }
  }
  const int ntile = (n + kSTile - 1) / kSTile;
  const int ntri = ntile * (ntile + 1) / 2;
  const int nT = N / (2 * lv.s);
  int nF = 0;  // fill-in tasks (k_node32 forms the fill-in blocks itself)
  if (!fused)
    for (int t = 1; t < lv.dim; ++t)
      if (lv.s - 1 + 2 * lv.s * t + lv.s < N) ++nF;
  const int tasks = nT * ntri + nF * ntile * ntile;
  if (tasks > 0) {
    kt_begin(kKcSchur, st);
    if (xf != nullptr || use_schur_small(B * (nT + nF), n, (int)sizeof(T))) {
      const int nsmall = nT * (NB * (NB + 1) / 2) + nF * NB * NB;
      const size_t lds_small = schur_small_lds<T, NB>();
      auto k = k_schur_small<T, NB, NF, false>;
      if constexpr (sizeof(T) == 4)
        if (bf16) k = k_schur_small<T, NB, NF, true>;
      hipLaunchKernelGGL(k, dim3(nsmall, B), dim3(kThreads), lds_small, st, L, E, mask, N, n, nE, lv.s, lv.off,
                         lv.next_off, nT, Lsrc, xf, m);
    } else {
      auto k = k_schur_blk<T, NF, false>;
      if constexpr (sizeof(T) == 4)
        if (bf16) k = k_schur_blk<T, NF, true>;
      hipLaunchKernelGGL(k, dim3(8 * schur_blk_per_xcd(nT, nF, ntile), B), dim3(kThreads), 0, st, L, E, mask, N, n, nE,
                         lv.s, lv.off, lv.next_off, nT, nF, ntile, Lsrc);
    }
    kt_end(kKcSchur, st);
  }
  return hipGetLastError();
}

// One factor level for kMaxN < n <= kMaxNLarge: the tiled potrf and trsm of
// nd_large.hip, then the Schur updates of k_schur_blk (which walks any n).
template <typename T>
static hipError_t factor_level_large(T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, hipStream_t st,
                                     hipEvent_t after_potrf, hipEvent_t after_trsm, hipEvent_t before_trsm) {
  const int nE = (int)off_diag_storage_len(N);
  const int Lv = recursive_iterations(N);
  const Level lv = level_info(N, lidx);
  if (lv.dim <= 0 || B <= 0) return hipSuccess;
  kt_begin(kKcPotrf, st);
  hipError_t e = potrf_large<T>(L, mask, B, N, n, lidx, st);
  kt_end(kKcPotrf, st);
  if (e == hipSuccess && after_potrf != nullptr) e = hipEventRecord(after_potrf, st);
  if (e != hipSuccess || lidx == Lv - 1) return e;
  if (before_trsm != nullptr && (e = hipStreamWaitEvent(st, before_trsm, 0)) != hipSuccess) return e;
  kt_begin(kKcTrsm, st);
  e = trsm_large<T>(L, E, mask, B, N, n, lidx, st);
  kt_end(kKcTrsm, st);
  if (e == hipSuccess && after_trsm != nullptr) e = hipEventRecord(after_trsm, st);
  if (e != hipSuccess) return e;
  const int ntile = (n + kSTile - 1) / kSTile;
  const int ntri = ntile * (ntile + 1) / 2;
  const int nT = N / (2 * lv.s);
  int nF = 0;
  for (int t = 1; t < lv.dim; ++t)
    if (lv.s - 1 + 2 * lv.s * t + lv.s < N) ++nF;
  const int tasks = nT * ntri + nF * ntile * ntile;
  if (tasks > 0) {
    kt_begin(kKcSchur, st);
    hipLaunchKernelGGL((k_schur_blk<T, 0, false>), dim3(8 * schur_blk_per_xcd(nT, nF, ntile), B), dim3(kThreads), 0, st,
                       L, E, mask, N, n, nE, lv.s, lv.off, lv.next_off, nT, nF, ntile, (const T*)nullptr);
    kt_end(kKcSchur, st);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t factor_level_blocked(T* L, T* E, const uint8_t* mask, int B, int N, int n, int lidx, int prec,
                                hipStream_t st, hipEvent_t after_potrf, hipEvent_t after_trsm,
                                hipEvent_t before_trsm, const T* Lsrc, const T* Esrc) {
  if (n > kMaxN) {
    if (Lsrc != nullptr || Esrc != nullptr) return hipErrorInvalidValue;  // in place only
    return factor_level_large<T>(L, E, mask, B, N, n, lidx, st, after_potrf, after_trsm, before_trsm);
  }
  const int fp = prec;
  switch (n) {
    case 32: return factor_level_nb<T, 1, 32>(L, E, mask, B, N, n, lidx, fp, st, after_potrf, after_trsm, before_trsm, Lsrc,
                                                   Esrc);
    case 64: return factor_level_nb<T, 2, 64>(L, E, mask, B, N, n, lidx, fp, st, after_potrf, after_trsm, before_trsm, Lsrc,
                                                   Esrc);
    case 96: return factor_level_nb<T, 3, 96>(L, E, mask, B, N, n, lidx, fp, st, after_potrf, after_trsm, before_trsm, Lsrc,
                                                   Esrc);
    case 128: return factor_level_nb<T, 4, 128>(L, E, mask, B, N, n, lidx, fp, st, after_potrf, after_trsm, before_trsm, Lsrc,
                                                   Esrc);
    default: break;
  }
  switch ((n + 31) / 32) {
    case 1: return factor_level_nb<T, 1, 0>(L, E, mask, B, N, n, lidx, fp, st, after_potrf, after_trsm, before_trsm, Lsrc,
                                                   Esrc);
    case 2: return factor_level_nb<T, 2, 0>(L, E, mask, B, N, n, lidx, fp, st, after_potrf, after_trsm, before_trsm, Lsrc,
                                                   Esrc);
    case 3: return factor_level_nb<T, 3, 0>(L, E, mask, B, N, n, lidx, fp, st, after_potrf, after_trsm, before_trsm, Lsrc,
                                                   Esrc);
    case 4: return factor_level_nb<T, 4, 0>(L, E, mask, B, N, n, lidx, fp, st, after_potrf, after_trsm, before_trsm, Lsrc,
                                                   Esrc);
    default: return hipErrorInvalidValue;
  }
}

template <typename T, int NB, int NF>
static hipError_t trtri_nb(const T* L, T* W, const uint8_t* mask, int B, int N, int n, hipStream_t st) {
  if (B <= 0) return hipSuccess;
  const size_t lds = trtri_blk_lds<T, NB>();
  kt_begin(kKcForward, st);
  hipLaunchKernelGGL((k_trtri_blk<T, NB, NF>), dim3(N, B), dim3(64 * NB), lds, st, L, W, mask, N, n);
  kt_end(kKcForward, st);
  return hipGetLastError();
}

template <typename T>
hipError_t trtri_blocked(const T* L, T* W, const uint8_t* mask, int B, int N, int n, hipStream_t st) {
  if (n > kMaxN) {
    kt_begin(kKcForward, st);
    const hipError_t e = trtri_large<T>(L, W, mask, B, N, n, st);
    kt_end(kKcForward, st);
    return e;
  }
  switch (n) {
    case 32: return trtri_nb<T, 1, 32>(L, W, mask, B, N, n, st);
    case 64: return trtri_nb<T, 2, 64>(L, W, mask, B, N, n, st);
    case 96: return trtri_nb<T, 3, 96>(L, W, mask, B, N, n, st);
    case 128: return trtri_nb<T, 4, 128>(L, W, mask, B, N, n, st);
    default: break;
  }
  switch ((n + 31) / 32) {
    case 1: return trtri_nb<T, 1, 0>(L, W, mask, B, N, n, st);
    case 2: return trtri_nb<T, 2, 0>(L, W, mask, B, N, n, st);
    case 3: return trtri_nb<T, 3, 0>(L, W, mask, B, N, n, st);
    case 4: return trtri_nb<T, 4, 0>(L, W, mask, B, N, n, st);
    default: return hipErrorInvalidValue;
  }
}

template <typename T, int NB, int NF>
static hipError_t forward_step_nb(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                  int step, hipStream_t st, bool apply_next, bool update = true) {
  const int nE = (int)off_diag_storage_len(N);
  const int mch = (m + kRhsChunk - 1) / kRhsChunk;
  if (B <= 0 || mch <= 0) return hipSuccess;
  kt_begin(kKcForward, st);
  // step l's updates and level l+1's applies in one launch (k_sweep32,
  // k_sweep_nb)
  if (update && step >= 0) {
    const Level lv = level_info(N, step);
    const int items = N / (2 * lv.s);
    if (items > 0) {
      const int ap = (apply_next && level_info(N, step + 1).dim > 0) ? 1 : 0;
      if constexpr (NF == 32) {
        const size_t lds = sweep32_lds<T>(m);
        hipLaunchKernelGGL((k_sweep32<T, true>), dim3(items, mch, B), dim3(kThreads), lds, st, L, E, x, mask, N, m,
                           nE, lv.s, lv.off, ap);
      } else {
        const size_t lds = sweepnb_lds<T, NB>(m);
        hipLaunchKernelGGL((k_sweep_nb<T, NB, NF, true>), dim3(items, mch, B), dim3(sweepnb_threads<T, NB, true>()), lds,
                           st, L, E, x, mask, N, n, m, nE, lv.s, lv.off, ap);
      }
    }
    kt_end(kKcForward, st);
    return hipGetLastError();
  }
  const Level nx = level_info(N, step + 1);
  if (apply_next && nx.dim > 0) {
    const size_t lds = apply_fwd_lds<T, NB>(m);
    hipLaunchKernelGGL((k_apply_fwd<T, NB, NF>), dim3(nx.dim, mch, B), dim3(kThreads), lds, st, L, x, mask, N, n, m,
                       nx.s);
  }
  kt_end(kKcForward, st);
  return hipGetLastError();
}

template <typename T>
static hipError_t forward_step_parts(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                              int step, hipStream_t st, bool a, bool u) {
  if (n > kMaxN) {
    hipError_t e = hipSuccess;
    kt_begin(kKcForward, st);
    if (u && step >= 0) e = update_large<T>(E, x, mask, B, N, n, m, step, false, st);
    if (e == hipSuccess && a) e = apply_large<T>(L, x, mask, B, N, n, m, step + 1, false, st);
    kt_end(kKcForward, st);
    return e;
  }
  switch (n) {
    case 32: return forward_step_nb<T, 1, 32>(L, E, x, mask, B, N, n, m, step, st, a, u);
    case 64: return forward_step_nb<T, 2, 64>(L, E, x, mask, B, N, n, m, step, st, a, u);
    case 96: return forward_step_nb<T, 3, 96>(L, E, x, mask, B, N, n, m, step, st, a, u);
    case 128: return forward_step_nb<T, 4, 128>(L, E, x, mask, B, N, n, m, step, st, a, u);
    default: break;
  }
  switch ((n + 31) / 32) {
    case 1: return forward_step_nb<T, 1, 0>(L, E, x, mask, B, N, n, m, step, st, a, u);
    case 2: return forward_step_nb<T, 2, 0>(L, E, x, mask, B, N, n, m, step, st, a, u);
    case 3: return forward_step_nb<T, 3, 0>(L, E, x, mask, B, N, n, m, step, st, a, u);
    case 4: return forward_step_nb<T, 4, 0>(L, E, x, mask, B, N, n, m, step, st, a, u);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t forward_step_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                int step, hipStream_t st, bool apply_next) {
  return forward_step_parts<T>(L, E, x, mask, B, N, n, m, step, st, apply_next, true);
}

template <typename T>
hipError_t forward_update_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                  int step, hipStream_t st) {
  return forward_step_parts<T>(L, E, x, mask, B, N, n, m, step, st, false, true);
}

template <typename T>
hipError_t forward_apply_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                 int level, hipStream_t st) {
  return forward_step_parts<T>(L, E, x, mask, B, N, n, m, level - 1, st, true, false);
}

template <typename T, int NB, int NF>
static hipError_t backward_level_nb(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                    int lidx, hipStream_t st) {
  const int nE = (int)off_diag_storage_len(N);
  const int mch = (m + kRhsChunk - 1) / kRhsChunk;
  const Level lv = level_info(N, lidx);
  if (lv.dim <= 0 || B <= 0 || mch <= 0) return hipSuccess;
  kt_begin(kKcBackward, st);
  if constexpr (NF == 32) {  // update and apply in one launch
    const size_t lds = sweep32_lds<T>(m);
    hipLaunchKernelGGL((k_sweep32<T, false>), dim3(lv.dim, mch, B), dim3(kThreads), lds, st, L, E, x, mask, N, m, nE,
                       lv.s, lv.off, 1);
    kt_end(kKcBackward, st);
    return hipGetLastError();
  } else {
    const size_t lds = sweepnb_lds<T, NB>(m);
    hipLaunchKernelGGL((k_sweep_nb<T, NB, NF, false>), dim3(lv.dim, mch, B), dim3(sweepnb_threads<T, NB, false>()), lds,
                       st, L, E, x, mask, N, n, m, nE, lv.s, lv.off, 1);
    kt_end(kKcBackward, st);
    return hipGetLastError();
  }
}

template <typename T>
hipError_t backward_level_blocked(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m,
                                  int lidx, hipStream_t st) {
  if (n > kMaxN) {
    const Level lv = level_info(N, lidx);
    hipError_t e = hipSuccess;
    kt_begin(kKcBackward, st);
    if (!(lv.dim == 1 && lv.s - 1 + lv.s >= N)) e = update_large<T>(E, x, mask, B, N, n, m, lidx, true, st);
    if (e == hipSuccess) e = apply_large<T>(L, x, mask, B, N, n, m, lidx, true, st);
    kt_end(kKcBackward, st);
    return e;
  }
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

hipError_t overlap_create(SweepOverlap& ov) {
  hipError_t e = hipStreamCreateWithFlags(&ov.side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ov.fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ov.join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ov.ready, hipEventDisableTiming);
  for (int l = 0; l < kSweepMaxLevels && e == hipSuccess; ++l) {
    e = hipEventCreateWithFlags(&ov.potrf[l], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ov.trsm[l], hipEventDisableTiming);
  }
  if (e != hipSuccess) overlap_destroy(ov);
  return e;
}

void overlap_destroy(SweepOverlap& ov) {
  for (int l = 0; l < kSweepMaxLevels; ++l) {
    if (ov.potrf[l]) (void)hipEventDestroy(ov.potrf[l]);
    if (ov.trsm[l]) (void)hipEventDestroy(ov.trsm[l]);
  }
  if (ov.fork) (void)hipEventDestroy(ov.fork);
  if (ov.join) (void)hipEventDestroy(ov.join);
  if (ov.ready) (void)hipEventDestroy(ov.ready);
  if (ov.side) (void)hipStreamDestroy(ov.side);
  ov = SweepOverlap();
}

template <typename T>
bool fused_forward_applies(int n, int prec) {
  return n == 32 && !(sizeof(T) == 4 && (prec & (kPrecF64Pivots | kPrecBf16Schur)) != 0);
}

template <typename T>
hipError_t factor_and_forward(T* L, T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, int prec,
                              hipStream_t st, const T* Lsrc, const T* Esrc) {
  if (!fused_forward_applies<T>(n, prec)) return hipErrorInvalidValue;
  const int Lv = recursive_iterations(N);
  hipError_t e = hipSuccess;
  for (int l = 0; l < Lv && e == hipSuccess; ++l)
    e = factor_level_nb<T, 1, 32>(L, E, mask, B, N, n, l, prec, st, nullptr, nullptr, nullptr, l == 0 ? Lsrc : nullptr,
                                  l == 0 ? Esrc : nullptr, x, m);
  return e;
}

// The reference's fused factor_and_solve likewise interleaves the forward
// sweep into the factor's levels (block_tridiag_solver.py:926-1233).  Here
// the sweep runs on a second stream: the inverses of level l follow that
// level's potrf, its updates into the receiving blocks follow its trsm.
// Both read only finished factor blocks and write only x, which the factor
// does not touch.
template <typename T>
hipError_t factor_with_forward(T* L, T* E, T* x, const uint8_t* factor_mask, const uint8_t* sweep_mask, int B, int N,
                               int n, int m, int prec, const SweepOverlap& ov, hipStream_t st,
                               const std::function<hipError_t(hipStream_t)>* side_first, const T* Lsrc,
                               const T* Esrc) {
  // n == 32: the forward sweep runs inside the factor's own kernels
  if (side_first == nullptr && factor_mask == sweep_mask && fused_forward_applies<T>(n, prec))
    return factor_and_forward<T>(L, E, x, factor_mask, B, N, n, m, prec, st, Lsrc, Esrc);
  const int Lv = recursive_iterations(N);
  hipError_t e = hipEventRecord(ov.fork, st);
  if (e == hipSuccess) e = hipStreamWaitEvent(ov.side, ov.fork, 0);
  if (e == hipSuccess && side_first != nullptr) e = (*side_first)(ov.side);
  if (e == hipSuccess && side_first != nullptr) e = hipEventRecord(ov.ready, ov.side);
  for (int l = 0; l < Lv && e == hipSuccess; ++l) {
    const bool last = l == Lv - 1;
    e = factor_level_blocked<T>(L, E, factor_mask, B, N, n, l, prec, st, ov.potrf[l],
                                last ? nullptr : ov.trsm[l], (l == 0 && side_first != nullptr) ? ov.ready : nullptr,
                                l == 0 ? Lsrc : nullptr, l == 0 ? Esrc : nullptr);
    if (e == hipSuccess) e = hipStreamWaitEvent(ov.side, ov.potrf[l], 0);
    if (e == hipSuccess) e = forward_apply_blocked<T>(L, E, x, sweep_mask, B, N, n, m, l, ov.side);
    if (e == hipSuccess && !last) e = hipStreamWaitEvent(ov.side, ov.trsm[l], 0);
    if (e == hipSuccess && !last) e = forward_update_blocked<T>(L, E, x, sweep_mask, B, N, n, m, l, ov.side);
  }
  if (e == hipSuccess) e = hipEventRecord(ov.join, ov.side);
  if (e == hipSuccess) e = hipStreamWaitEvent(st, ov.join, 0);
  return e;
}

template bool fused_forward_applies<float>(int, int);
template bool fused_forward_applies<double>(int, int);
template hipError_t factor_and_forward<float>(float*, float*, float*, const uint8_t*, int, int, int, int, int,
                                              hipStream_t, const float*, const float*);
template hipError_t factor_and_forward<double>(double*, double*, double*, const uint8_t*, int, int, int, int, int,
                                               hipStream_t, const double*, const double*);
template hipError_t factor_with_forward<float>(float*, float*, float*, const uint8_t*, const uint8_t*, int, int, int,
                                               int, int, const SweepOverlap&, hipStream_t,
                                               const std::function<hipError_t(hipStream_t)>*, const float*,
                                               const float*);
template hipError_t factor_with_forward<double>(double*, double*, double*, const uint8_t*, const uint8_t*, int, int,
                                                int, int, int, const SweepOverlap&, hipStream_t,
                                                const std::function<hipError_t(hipStream_t)>*, const double*,
                                                const double*);

template hipError_t trtri_blocked<float>(const float*, float*, const uint8_t*, int, int, int, hipStream_t);
template hipError_t trtri_blocked<double>(const double*, double*, const uint8_t*, int, int, int, hipStream_t);
template hipError_t factor_level_blocked<float>(float*, float*, const uint8_t*, int, int, int, int, int,
                                                hipStream_t, hipEvent_t, hipEvent_t, hipEvent_t, const float*,
                                                const float*);
template hipError_t factor_level_blocked<double>(double*, double*, const uint8_t*, int, int, int, int, int,
                                                 hipStream_t, hipEvent_t, hipEvent_t, hipEvent_t, const double*,
                                                 const double*);
template hipError_t forward_step_blocked<float>(const float*, const float*, float*, const uint8_t*, int, int, int, int,
                                                int, hipStream_t, bool);
template hipError_t forward_step_blocked<double>(const double*, const double*, double*, const uint8_t*, int, int, int,
                                                 int, int, hipStream_t, bool);
template hipError_t forward_update_blocked<float>(const float*, const float*, float*, const uint8_t*, int, int, int,
                                                  int, int, hipStream_t);
template hipError_t forward_update_blocked<double>(const double*, const double*, double*, const uint8_t*, int, int,
                                                   int, int, int, hipStream_t);
template hipError_t forward_apply_blocked<float>(const float*, const float*, float*, const uint8_t*, int, int, int,
                                                 int, int, hipStream_t);
template hipError_t forward_apply_blocked<double>(const double*, const double*, double*, const uint8_t*, int, int,
                                                  int, int, int, hipStream_t);
template hipError_t backward_level_blocked<float>(const float*, const float*, float*, const uint8_t*, int, int, int,
                                                  int, int, hipStream_t);
template hipError_t backward_level_blocked<double>(const double*, const double*, double*, const uint8_t*, int, int,
                                                   int, int, int, hipStream_t);

}  // namespace socu
