// Wavefront-level building blocks for the blocked ND kernels on gfx950.
//
// * MFMA 16x16x4 wrappers for fp64 and fp32 with their lane layouts:
//     A (16x4):  lane l holds A[l%16][l/16]
//     B (4x16):  lane l holds B[l/16][l%16]
//     D (16x16): fp32: lane l, reg r holds D[4*(l/16)+r][l%16]
//                fp64: lane l, reg r holds D[4*r+(l/16)][l%16]
// * wave_gemm: a 32x32 (2x2 MFMA tiles) accumulation C += A B with A, B
//   read from LDS through element strides.
// * tile_chol_inv: Cholesky and inverse of a 32x32 diagonal tile by one
//   wavefront, row r in lane r's registers, column broadcasts through LDS.
#pragma once

#include <hip/hip_runtime.h>

#include "nd_common.h"

namespace socu {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef f64x4 acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * r + (l >> 4); }
  static __device__ __forceinline__ int col(int l, int r) { return l & 15; }
};

template <>
struct Mfma<float> {
  typedef f32x4 acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
  static __device__ __forceinline__ int col(int l, int r) { return l & 15; }
};

template <typename T>
__device__ __forceinline__ typename Mfma<T>::acc_t acc_zero() {
  typename Mfma<T>::acc_t z;
  z[0] = z[1] = z[2] = z[3] = T(0);
  return z;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// Value of `v` in lane `src` (src must be wave-uniform).
__device__ __forceinline__ double bcast_lane(double v, int src) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_readlane(lo, src);
  hi = __builtin_amdgcn_readlane(hi, src);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float bcast_lane(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// C(32x32) += A(32xK) B(Kx32); A(i,k) = A0[i*ai + k*ak], B(k,j) = B0[k*bk + j*bj].
// K must be a multiple of 4.  acc[ti][tj] is the 16x16 MFMA tile (ti, tj).
template <typename T>
__device__ __forceinline__ void wave_gemm(typename Mfma<T>::acc_t acc[2][2], const T* A0, int ai, int ak, const T* B0,
                                          int bk, int bj, int K) {
  const int l = lane_id();
  const int i = l & 15, kq = l >> 4;
  for (int k = 0; k < K; k += 4) {
    const int kk = k + kq;
    const T a0 = A0[i * ai + kk * ak];
    const T a1 = A0[(i + 16) * ai + kk * ak];
    const T b0 = B0[kk * bk + i * bj];
    const T b1 = B0[kk * bk + (i + 16) * bj];
    acc[0][0] = Mfma<T>::mma(a0, b0, acc[0][0]);
    acc[0][1] = Mfma<T>::mma(a0, b1, acc[0][1]);
    acc[1][0] = Mfma<T>::mma(a1, b0, acc[1][0]);
    acc[1][1] = Mfma<T>::mma(a1, b1, acc[1][1]);
  }
}

// C(32x32) += A(32xK) B(Kx32) with element fetch functors fa(i, k), fb(k, j)
// (for operands with implicit structure, e.g. triangular tiles).
template <typename T, typename FA, typename FB>
__device__ __forceinline__ void wave_gemm_f(typename Mfma<T>::acc_t acc[2][2], FA fa, FB fb, int K) {
  const int l = lane_id();
  const int i = l & 15, kq = l >> 4;
  for (int k = 0; k < K; k += 4) {
    const int kk = k + kq;
    const T a0 = fa(i, kk), a1 = fa(i + 16, kk);
    const T b0 = fb(kk, i), b1 = fb(kk, i + 16);
    acc[0][0] = Mfma<T>::mma(a0, b0, acc[0][0]);
    acc[0][1] = Mfma<T>::mma(a0, b1, acc[0][1]);
    acc[1][0] = Mfma<T>::mma(a1, b0, acc[1][0]);
    acc[1][1] = Mfma<T>::mma(a1, b1, acc[1][1]);
  }
}

// One 16x16 tile: C += A(16xK) B(Kx16), same stride convention.
template <typename T>
__device__ __forceinline__ void wave_gemm16(typename Mfma<T>::acc_t& acc, const T* A0, int ai, int ak, const T* B0,
                                            int bk, int bj, int K) {
  const int l = lane_id();
  const int i = l & 15, kq = l >> 4;
  for (int k = 0; k < K; k += 4) {
    const int kk = k + kq;
    acc = Mfma<T>::mma(A0[i * ai + kk * ak], B0[kk * bk + i * bj], acc);
  }
}

// BF16 products with FP32 accumulation (v_mfma_f32_16x16x16_bf16): lane l
// holds A[l%16][4(l/16) .. 4(l/16)+3] and B[4(l/16) .. 4(l/16)+3][l%16]
// (the k order inside a group is the same for A and B, which is all the sum
// needs); D has the FP32 16x16x4 layout.  Operands come from FP32 arrays and
// are rounded to BF16 (nearest, ties to even) as they are read.
typedef short bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f32_to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);  // finite inputs: round to nearest even
  return (short)(u >> 16);
}

__device__ __forceinline__ f32x4 mma_bf16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// As wave_gemm<float> with the products on BF16 MFMA; K a multiple of 16.
__device__ __forceinline__ void wave_gemm_bf16(f32x4 acc[2][2], const float* A0, int ai, int ak, const float* B0,
                                               int bk, int bj, int K) {
  const int l = lane_id();
  const int i = l & 15, kq = 4 * (l >> 4);
  for (int k = 0; k < K; k += 16) {
    bf16x4 a0, a1, b0, b1;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int kk = k + kq + t;
      a0[t] = f32_to_bf16(A0[i * ai + kk * ak]);
      a1[t] = f32_to_bf16(A0[(i + 16) * ai + kk * ak]);
      b0[t] = f32_to_bf16(B0[kk * bk + i * bj]);
      b1[t] = f32_to_bf16(B0[kk * bk + (i + 16) * bj]);
    }
    acc[0][0] = mma_bf16(a0, b0, acc[0][0]);
    acc[0][1] = mma_bf16(a0, b1, acc[0][1]);
    acc[1][0] = mma_bf16(a1, b0, acc[1][0]);
    acc[1][1] = mma_bf16(a1, b1, acc[1][1]);
  }
}

// As wave_gemm16<float> with the products on BF16 MFMA; K a multiple of 16.
__device__ __forceinline__ void wave_gemm16_bf16(f32x4& acc, const float* A0, int ai, int ak, const float* B0, int bk,
                                                 int bj, int K) {
  const int l = lane_id();
  const int i = l & 15, kq = 4 * (l >> 4);
  for (int k = 0; k < K; k += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int kk = k + kq + t;
      a[t] = f32_to_bf16(A0[i * ai + kk * ak]);
      b[t] = f32_to_bf16(B0[kk * bk + i * bj]);
    }
    acc = mma_bf16(a, b, acc);
  }
}

// Schur-update products in the storage precision, or (BF, float storage) on
// BF16 MFMA.
template <typename T, bool BF>
__device__ __forceinline__ void schur_gemm(typename Mfma<T>::acc_t acc[2][2], const T* A0, int ai, int ak,
                                           const T* B0, int bk, int bj, int K) {
  if constexpr (BF) wave_gemm_bf16(acc, A0, ai, ak, B0, bk, bj, K);
  else wave_gemm<T>(acc, A0, ai, ak, B0, bk, bj, K);
}
template <typename T, bool BF>
__device__ __forceinline__ void schur_gemm16(typename Mfma<T>::acc_t& acc, const T* A0, int ai, int ak, const T* B0,
                                             int bk, int bj, int K) {
  if constexpr (BF) wave_gemm16_bf16(acc, A0, ai, ak, B0, bk, bj, K);
  else wave_gemm16<T>(acc, A0, ai, ak, B0, bk, bj, K);
}

// Apply f(row, col, value) to every element of a 32x32 accumulator.
template <typename T, typename F>
__device__ __forceinline__ void acc_foreach(typename Mfma<T>::acc_t acc[2][2], F f) {
  const int l = lane_id();
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) f(16 * ti + Mfma<T>::row(l, r), 16 * tj + Mfma<T>::col(l, r), acc[ti][tj][r]);
}

template <typename T, typename F>
__device__ __forceinline__ void acc16_foreach(typename Mfma<T>::acc_t& acc, F f) {
  const int l = lane_id();
#pragma unroll
  for (int r = 0; r < 4; ++r) f(Mfma<T>::row(l, r), Mfma<T>::col(l, r), acc[r]);
}

// Cholesky and inverse of a 32x32 tile, executed by one whole wavefront.
// S: tile in LDS, row-major with `pitch`; its lower triangle is the input.
// On return the lower triangle holds L, the strict upper triangle holds the
// transposed inverse (S[c][q] = Linv[q][c] for q > c) and dinv[r] = 1/L[r][r].
// col: 128 entries of wave-private LDS (two buffers of two columns; its
// first 64 entries are overwritten at the end).
// ut: a 32 x 33 LDS tile that receives Linv transposed in full (ut[r][c] =
// Linv[c][r], zeros and diagonal included).
//
// Lanes 0-31 keep row r of the tile in u[]; lanes 32-63 keep column r of the
// inverse (Gauss-Jordan: Linv = L_31^-1 ... L_0^-1).  The pivots are taken two
// at a time: for the 2x2 pivot block [a b; b c] (from lanes k, k+1 by
// readlane), 1/sqrt(a) and 1/sqrt(ac - b^2) are computed side by side and
// 1/sqrt(c - b^2/a) = sqrt(a) / sqrt(ac - b^2) follows with two products, so
// the serial chain per two columns is one reciprocal square root.  Steps k
// and k+1 then fuse into one rank-2 update per entry,
//   u[q] -= (c1 - c2 t) A[q][k] + c2 A[q][k+1],   t = b / a,
// with each half's coefficients c1 (step k) and c2 (step k+1, from the entry
// k+1 updated by step k).  Columns k, k+1 come from the LDS buffer (k/2) % 2
// (LDS operations of one wavefront complete in issue order, so no barrier is
// needed).  Look-ahead: entries k+2, k+3 are updated first and published as
// the next pair's columns.  FP32 (EARLY): the next pair's column reads are
// issued right after (a scheduling barrier keeps them there), so that their
// LDS latency runs under this pair's remaining updates instead of the next
// pivot chain; FP64 reads them at the start of the pair step (two sets of FP64
// columns in registers spill).
// The coefficients are not masked for row lanes at or above the pair: those
// lanes only change entries of their strict upper triangle, which nothing
// reads (not the column buffers, the pivots, nor the store), and their own
// pivot entries are replaced explicitly.
// hook(k) runs at the start of pair step k (k even, a compile-time constant
// after unrolling): a caller can issue loads there, late in the chain, when
// the column registers of the early steps are free.
struct NoHook {
  __device__ __forceinline__ void operator()(int) const {}
};
template <typename T, typename Hook = NoHook>
__device__ __forceinline__ void tile_chol_inv(T* S, int pitch, T* dinv, T* col, T* ut, Hook hook = Hook()) {
  const int lane = lane_id();
  const int r = lane & 31;
  const bool rowlane = lane < 32;
  T u[32];
  {
    // Row lanes take row r as it is, strict upper entries included: those
    // entries only ever meet the multipliers of their own lane (see below)
    // and are not stored.  The reads are unconditional (in bounds) and
    // pinned, so that they are all in flight at once; inverse lanes then
    // replace theirs by the unit vector e_r.  (Selects on the reads let the
    // compiler turn every read into a branch with a wait of its own.)
#pragma unroll
    for (int c = 0; c < 32; ++c) u[c] = S[r * pitch + c];
#pragma unroll
    for (int c = 0; c < 32; ++c) asm volatile("" : "+v"(u[c]));
    if (!rowlane) {
#pragma unroll
      for (int c = 0; c < 32; ++c) u[c] = (c == r) ? T(1) : T(0);
    }
  }
  // one base address in a vector register: every column access below is
  // that base plus an immediate offset
  typedef __attribute__((address_space(3))) T lds_t;
  lds_t* colv = (lds_t*)col;
  asm volatile("" : "+v"(colv));
  if (rowlane) {
    colv[r] = u[0];
    colv[32 + r] = u[1];
  }
  // Publishing a pair's columns: row lanes write buffer p, inverse lanes the
  // other buffer, whose columns are already in registers (so no branch on
  // the lane kind): pubv[p] is each lane's target for buffer p.
  lds_t* pubv0 = colv + (rowlane ? 0 : 64);
  lds_t* pubv1 = colv + (rowlane ? 64 : 0);
  asm volatile("" : "+v"(pubv0), "+v"(pubv1));
  // Row index for the pivot selects below; inverse lanes take one past every
  // pivot, so that the selects give them the multipliers (no branch).
  const int rsel = rowlane ? r : 64;
  constexpr bool EARLY = sizeof(T) == 4;
  // FP32 column reads are 16-byte vector reads of four entries from an
  // aligned start (col is 16-byte aligned; entries below the first one
  // needed are dropped): half the LDS instructions of paired 4-byte reads.
  typedef T t4 __attribute__((ext_vector_type(4)));
  typedef __attribute__((address_space(3))) t4 lds_t4;
  T c0[32], c1[32];
  if constexpr (EARLY) {
#pragma unroll
    for (int q0 = 0; q0 < 32; q0 += 4) {
      const t4 v0 = *(const lds_t4*)(colv + q0);
      const t4 v1 = *(const lds_t4*)(colv + 32 + q0);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (q0 + j >= 2) {
          c0[q0 + j] = v0[j];
          c1[q0 + j] = v1[j];
        }
    }
  }
#pragma unroll
  for (int k = 0; k < 32; k += 2) {
    lds_t* nb = colv + 64 * (((k >> 1) + 1) & 1);
    lds_t* pb = (((k >> 1) + 1) & 1) ? pubv1 : pubv0;
    if constexpr (!EARLY) {
      const lds_t* cb = colv + 64 * ((k >> 1) & 1);
#pragma unroll
      for (int q = k + 2; q < 32; ++q) {
        c0[q] = cb[q];
        c1[q] = cb[32 + q];
      }
    }
    hook(k);
    const T a = bcast_lane(u[k], k);
    const T b = bcast_lane(u[k], k + 1);
    const T c = bcast_lane(u[k + 1], k + 1);
    const T inv1 = rsqrt_t(a);
    const T g = rsqrt_t(a * c - b * b);
    const T inv2 = g * (a * inv1);  // 1 / sqrt(c - b^2 / a)
    const T t = b * inv1 * inv1;
    // step k: row lanes l = L[r][k]; inverse lanes z_k
    const T lk1 = u[k] * inv1;
    const T cf1 = lk1 * inv1;
    // step k+1, on entry k+1 as updated by step k (A[k+1][k] = b)
    const T uk1 = u[k + 1] - cf1 * b;
    const T lk2 = uk1 * inv2;
    const T cf2 = lk2 * inv2;
    const T cf1p = cf1 - cf2 * t;
    // (explicit fused multiply-adds: two per entry and pair of steps)
    T ncf1 = -cf1p, ncf2 = -cf2;
    T n0[32], n1[32];
    if (k + 2 < 32) {
      u[k + 2] = fma(ncf2, c1[k + 2], fma(ncf1, c0[k + 2], u[k + 2]));
      u[k + 3] = fma(ncf2, c1[k + 3], fma(ncf1, c0[k + 3], u[k + 3]));
      asm volatile("" : "+v"(u[k + 2]), "+v"(u[k + 3]));
      pb[r] = u[k + 2];
      pb[32 + r] = u[k + 3];
      // the remaining updates use the coefficients as pinned here, after the
      // stores
      asm volatile("" : "+v"(ncf1), "+v"(ncf2) : : "memory");
      if constexpr (EARLY) {
#pragma unroll
        for (int q0 = (k + 4) & ~3; q0 < 32; q0 += 4) {
          const t4 v0 = *(const lds_t4*)(nb + q0);
          const t4 v1 = *(const lds_t4*)(nb + 32 + q0);
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (q0 + j >= k + 4) {
              n0[q0 + j] = v0[j];
              n1[q0 + j] = v1[j];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int q = k + 4; q < 32; ++q) u[q] = fma(ncf2, c1[q], fma(ncf1, c0[q], u[q]));
    const T pk1 = a * inv1;                // sqrt(a)
    const T pk2 = (c - b * t) * inv2;      // sqrt(c - b^2 / a)
    // lane comparisons on a copy of r the compiler cannot see through, so
    // that they are made here and not hoisted out of the loop (32 hoisted
    // masks spill to lanes of a vector register)
    // (rows above k and k+1 take the multipliers too: for them these are
    // strict-upper entries, which nothing reads)
    int rk = rsel;
    asm volatile("" : "+v"(rk));
    const bool on1 = rk == k, on2 = rk == k + 1;
    u[k] = on1 ? pk1 : lk1;
    u[k + 1] = on2 ? pk2 : lk2;
    // Materialise the registers every step: otherwise the optimiser sinks
    // the updates to their last use and keeps every broadcast value alive.
#pragma unroll
    for (int q = 0; q < 32; ++q) asm volatile("" : "+v"(u[q]));
    if constexpr (EARLY) {
#pragma unroll
      for (int q = k + 4; q < 32; ++q) {
        c0[q] = n0[q];
        c1[q] = n1[q];
      }
    }
  }
  // Lanes r and r + 32 share row r of S: the row lane stores columns c <= r,
  // the inverse lane columns c > r, i.e. lane stores columns lo .. lo + span
  // (none for inverse lane 31: lo = 32).  Every lane stores every entry, the
  // ones outside its range into its own entry of col (free by now; distinct
  // entries, so no bank conflicts): one compare and one address select per
  // entry, no branches (the range is hidden from the optimiser, which would
  // otherwise branch on the lane kind per entry).
  {
    lds_t* srow = (lds_t*)S + r * pitch;
    lds_t* sink = colv + lane;
    int lo = rowlane ? 0 : r + 1, span = rowlane ? r : (r < 31 ? 30 - r : 0);
    asm volatile("" : "+v"(lo), "+v"(span));
#pragma unroll
    for (int c = 0; c < 32; ++c) {
      lds_t* d = ((unsigned)(c - lo) <= (unsigned)span) ? srow + c : sink;
      asm volatile("" : "+v"(d));
      *d = u[c];
    }
  }
  // ut[r][c] = Linv[c][r] in full (zeros above the diagonal, 1/L[r][r] on
  // it): the inverse lanes' whole columns, the operand of the panel product
  if (!rowlane)
#pragma unroll
    for (int c = 0; c < 32; ++c) ut[r * 33 + c] = u[c];
  // dinv from the diagonal of ut (this wavefront's LDS writes above complete
  // first): tracking 1/L[r][r] through the pair steps by lane selects kept
  // 32 lane masks alive to the end
  if (rowlane) dinv[r] = ut[r * 33 + r];
}

// Inverse of a 32x32 lower-triangular tile (no factorisation), by one
// wavefront: lane c (of 0-31) forms column c of the inverse by forward
// substitution, x_r = (delta_rc - sum_{k<r} D[r][k] x_k) / D[r][r], with
// rows of D read as LDS broadcasts and four partial sums per row; lanes 32-63
// repeat lanes 0-31 and store nothing.  Output convention of tile_chol_inv:
// the strict upper triangle receives the transposed inverse
// (S[c][q] = Dinv[q][c] for q > c), dinv[r] = 1/D[r][r]; the lower triangle
// (the input) is unchanged.
template <typename T>
__device__ __forceinline__ void tile_tri_inv(T* S, int pitch, T* dinv) {
  const int lane = lane_id();
  const int c = lane & 31;
  T x[32];
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
#pragma unroll
    for (int k = 0; k + 3 < r; k += 4) {
      s0 = fma(S[r * pitch + k], x[k], s0);
      s1 = fma(S[r * pitch + k + 1], x[k + 1], s1);
      s2 = fma(S[r * pitch + k + 2], x[k + 2], s2);
      s3 = fma(S[r * pitch + k + 3], x[k + 3], s3);
    }
#pragma unroll
    for (int k = r & ~3; k < r; ++k) s0 = fma(S[r * pitch + k], x[k], s0);
    const T dr = T(1) / S[r * pitch + r];
    x[r] = (((r == c) ? T(1) : T(0)) - ((s0 + s1) + (s2 + s3))) * dr;
  }
  if (lane < 32) {
#pragma unroll
    for (int q = 0; q < 32; ++q)
      if (q > c) S[c * pitch + q] = x[q];
    dinv[c] = x[c];
  }
}

// Cholesky and inverse of a 16x16 tile by lanes 0-31 of the wavefront (row
// lanes 0-15, inverse lanes 16-31; lanes 32-63 must not call it), with the
// recurrence and in/out convention of tile_chol_inv, one pivot per step.
// col: 32 entries of wave-private LDS (the column buffer; all 32 are
// overwritten at the end).  As in tile_chol_inv, row lanes keep their strict
// upper entries unmasked (only their own updates touch them and they are not
// stored), the lane comparisons are made per step on a copy of the row index
// the compiler cannot see through, and the loads and stores carry no
// per-entry branches.
template <typename T>
__device__ __forceinline__ void tile_chol_inv16(T* S, int pitch, T* dinv, T* col) {
  const int lane = lane_id();
  const int r = lane & 15;
  const bool rowlane = lane < 16;
  typedef __attribute__((address_space(3))) T lds_t;
  T u[16];
  T rinv = T(1);
#pragma unroll
  for (int c = 0; c < 16; ++c) u[c] = S[r * pitch + c];
#pragma unroll
  for (int c = 0; c < 16; ++c) asm volatile("" : "+v"(u[c]));
  if (!rowlane) {
#pragma unroll
    for (int c = 0; c < 16; ++c) u[c] = (c == r) ? T(1) : T(0);
  }
  const int rsel = rowlane ? r : 64;  // inverse lanes: past every pivot
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (rowlane) col[r] = u[k];
    const T p = bcast_lane(u[k], k);
    const T inv = rsqrt_t(p);
    const T lk = u[k] * inv;
    const T coef = lk * inv;
#pragma unroll
    for (int q = k + 1; q < 16; ++q) u[q] -= coef * col[q];
    int rk = rsel;
    asm volatile("" : "+v"(rk));
    u[k] = (rk == k) ? p * inv : lk;
    rinv = (rk == k) ? inv : rinv;
    asm volatile("" : "+v"(rinv));
#pragma unroll
    for (int q = 0; q < 16; ++q) asm volatile("" : "+v"(u[q]));
  }
  // row lane r stores columns c <= r, inverse lane r columns c > r
  // (lo .. lo + span; none for inverse lane 15); the rest go to col[lane]
  {
    lds_t* srow = (lds_t*)S + r * pitch;
    lds_t* sink = (lds_t*)col + lane;
    int lo = rowlane ? 0 : r + 1, span = rowlane ? r : (r < 15 ? 14 - r : 0);
    asm volatile("" : "+v"(lo), "+v"(span));
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      lds_t* d = ((unsigned)(c - lo) <= (unsigned)span) ? srow + c : sink;
      asm volatile("" : "+v"(d));
      *d = u[c];
    }
  }
  if (rowlane) dinv[r] = rinv;
}

// One 16x16 MFMA product over K = 16: acc = A B with fa(i, k), fb(k, j).
template <typename T, typename FA, typename FB>
__device__ __forceinline__ typename Mfma<T>::acc_t mfma16x16x16(FA fa, FB fb) {
  const int l = lane_id();
  const int i = l & 15, kq = l >> 4;
  typename Mfma<T>::acc_t acc = acc_zero<T>();
#pragma unroll
  for (int k = 0; k < 16; k += 4) acc = Mfma<T>::mma(fa(i, k + kq), fb(k + kq, i), acc);
  return acc;
}

// 32x32 Cholesky + inverse by one wavefront through 16x16 halves:
// A11 = L11 L11^T; L21 = A21 L11^-T; A22 - L21 L21^T = L22 L22^T;
// Linv21 = -Linv22 L21 Linv11.  Same in/out convention as tile_chol_inv.
template <typename T>
__device__ __forceinline__ void tile_chol_inv32b(T* S, int pitch, T* dinv, T* col) {
  const int l = lane_id();
  T* S11 = S;
  T* S21 = S + 16 * pitch;
  T* S22 = S + 16 * pitch + 16;
  T* S12 = S + 16;
  if (l < 32) tile_chol_inv16<T>(S11, pitch, dinv, col);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // Linv11[a][b] (a >= b) from the factored S11
  auto linv11 = [&](int a, int b) -> T { return (b < a) ? S11[b * pitch + a] : ((a == b) ? dinv[a] : T(0)); };
  // L21 = A21 Linv11^T
  auto acc = mfma16x16x16<T>([&](int i, int k) { return S21[i * pitch + k]; },
                             [&](int k, int j) { return linv11(j, k); });
  acc16_foreach<T>(acc, [&](int i, int j, T v) { S21[i * pitch + j] = v; });
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // A22 -= L21 L21^T
  acc = mfma16x16x16<T>([&](int i, int k) { return S21[i * pitch + k]; },
                        [&](int k, int j) { return S21[j * pitch + k]; });
  acc16_foreach<T>(acc, [&](int i, int j, T v) { S22[i * pitch + j] -= v; });
  // T1 = L21 Linv11 (independent of the A22 factor)
  auto t1 = mfma16x16x16<T>([&](int i, int k) { return S21[i * pitch + k]; },
                            [&](int k, int j) { return linv11(k, j); });
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (l < 32) tile_chol_inv16<T>(S22, pitch, dinv + 16, col);
  // park T1 in the upper-right block (row-major), then Linv21 = -Linv22 T1
  acc16_foreach<T>(t1, [&](int i, int j, T v) { S12[i * pitch + j] = v; });
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  auto linv22 = [&](int a, int b) -> T { return (b < a) ? S22[b * pitch + a] : ((a == b) ? dinv[16 + a] : T(0)); };
  acc = mfma16x16x16<T>([&](int i, int k) { return linv22(i, k); }, [&](int k, int j) { return S12[k * pitch + j]; });
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // Linv21[i][j] stored transposed: S[j][16 + i]
  acc16_foreach<T>(acc, [&](int i, int j, T v) { S12[j * pitch + i] = -v; });
}

// Cholesky and inverse of G = 64/(2*NT) independent NT x NT tiles per
// wavefront (NT = 8, 16 or 32): lanes [g*2NT, g*2NT+NT) keep the rows of tile
// g, lanes [g*2NT+NT, (g+1)*2NT) its inverse columns, with the same recurrence
// as tile_chol_inv16.  Tile g is S0 + g*stride (row-major, pitch `pitch`), its
// 2NT-entry column buffer col0 + g*stride (all 2NT entries are overwritten at
// the end) and its 1/L[r][r] output dinv0 + g*stride.  The pivot comes
// through the column buffer, since each group has its own source lane.
template <typename T, int NT>
__device__ __forceinline__ void tile_chol_inv_grouped(T* S0, T* col0, T* dinv0, int stride, int pitch) {
  typedef __attribute__((address_space(3))) T lds_t;
  const int lane = lane_id();
  const int g = lane / (2 * NT);
  const int lr = lane % (2 * NT);
  const bool rowlane = lr < NT;
  const int r = lr % NT;
  T* St = S0 + g * stride;
  T* cb = col0 + g * stride;
  T u[NT];
  T rinv = T(1);
#pragma unroll
  for (int c = 0; c < NT; ++c) u[c] = St[r * pitch + c];
#pragma unroll
  for (int c = 0; c < NT; ++c) asm volatile("" : "+v"(u[c]));
  if (!rowlane) {
#pragma unroll
    for (int c = 0; c < NT; ++c) u[c] = (c == r) ? T(1) : T(0);
  }
  const int rsel = rowlane ? r : 64;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    if (rowlane) cb[r] = u[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const T p = cb[k];
    const T inv = rsqrt_t(p);
    const T lk = u[k] * inv;
    const T coef = lk * inv;
#pragma unroll
    for (int q = k + 1; q < NT; ++q) u[q] -= coef * cb[q];
    int rk = rsel;
    asm volatile("" : "+v"(rk));
    u[k] = (rk == k) ? p * inv : lk;
    rinv = (rk == k) ? inv : rinv;
    asm volatile("" : "+v"(rinv));
#pragma unroll
    for (int q = 0; q < NT; ++q) asm volatile("" : "+v"(u[q]));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  }
  {
    lds_t* srow = (lds_t*)St + r * pitch;
    lds_t* sink = (lds_t*)cb + lr;
    int lo = rowlane ? 0 : r + 1, span = rowlane ? r : (r < NT - 1 ? NT - 2 - r : 0);
    asm volatile("" : "+v"(lo), "+v"(span));
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      lds_t* d = ((unsigned)(c - lo) <= (unsigned)span) ? srow + c : sink;
      asm volatile("" : "+v"(d));
      *d = u[c];
    }
  }
  if (rowlane) dinv0[g * stride + r] = rinv;
}

}  // namespace socu
