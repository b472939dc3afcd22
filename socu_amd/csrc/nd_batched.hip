// Batched ND factor / solve for many independent systems of uneven size
// (horizon N_p, block size n_p and right-hand sides m_p <= 32 per system),
// one workgroup per system.  The reference can only batch systems that share
// N and n (block_tridiag_solver.py:749-768); this path has no counterpart
// there but computes, per system, the reference's factor and solve in the
// reference's storage.
//
// Schedule.  The workgroup walks the levels of its system.  At level l
// (s = 2^l, h = s/2) one phase runs, in rounds of TPW concurrent tasks:
//   * node tasks, one per level-l node k.  For l > 0 the task first applies
//     what the level l-1 elimination sends to node k: the Schur update
//     L_k -= R_{k-h} R_{k-h}^T + Lc_{k+h}^T Lc_{k+h} and the two fill-ins
//     that become node k's couplings, Lc_k -= R_{k-h} Lc_{k-h} and
//     R_k -= R_{k+h} Lc_{k+h} (the syrk / gemm updates of level l-1 in
//     block_tridiag_solver.py:797-813).  Then it factors and inverts L_k and
//     forms R_k <- R_k L_k^-T, Lc_k <- L_k^-1 Lc_k and, for the forward
//     sweep, y_k <- L_k^-1 (y_k - R_{k-h} y_{k-h} - Lc_{k+h}^T y_{k+h});
//   * deferred target tasks, one per level l-1 Schur target j that is not a
//     level-l node (j = 2s(v+1)-1): L_j -= R_{j-h} R_{j-h}^T +
//     Lc_{j+h}^T Lc_{j+h} and the matching y_j update.
// No two tasks of a level write the same block and every task reads only
// blocks finished at an earlier level, so a level is one barrier, with no
// atomics.  The backward sweep walks the levels down, one task per node:
// x_i <- L_i^-T (x_i - R_i^T x_{i+s} - Lc_i x_{i-s}).
//
// Tiles are NT x (NT+1) in LDS, NT in {8, 16, 32}, zero padded beyond n (the
// diagonal block identity padded).  A workgroup has W = 1, 2 or 4 wavefronts and
// runs one or more tasks at once (BShape); block products run on MFMA
// (16x16 tiles, NT >= 16) or, for NT = 8, as 2x2 register blocks per lane.
// Small workgroups (W = 1) let several systems share a compute unit, one per
// SIMD; the tuner picks W per size bucket.
#include "../../include/socu_hip.h"
#include "nd_common.h"
#include "nd_internal.h"
#include "nd_tile.h"

namespace socu {

namespace {

template <typename T, int NT>
struct BSlot {
  T A[NT * (NT + 1)];  // diagonal block being factored
  T B[NT * (NT + 1)];
  T C[NT * (NT + 1)];
  T D[NT * (NT + 1)];
  T col[2 * NT];  // column buffer and store sink of the tile factors
  T dinv[NT];
};

// Workgroups of W wavefronts.  A task uses LPT = min(64 W, NT^2/4) lanes
// (at most four tile elements per lane; NT = 32: 64 lanes, sixteen elements)
// and the workgroup runs TPW tasks at once: NT = 32 and NT = 16: one task per
// wavefront; NT = 8: four tasks per wavefront.
template <int NT, int W>
struct BShape {
  static constexpr int TP = NT + 1;
  static constexpr int BT = 64 * W;                  // threads per workgroup
  // lanes a task needs at 4 outputs per lane; NT = 32: one wavefront per
  // task (16 per lane), so that wider workgroups run several tasks of the
  // system at once instead of spreading one task (measured slower)
  static constexpr int NAT = NT == 32 ? 64 : NT * NT / 4;
  static constexpr int LPT = NAT < BT ? NAT : BT;    // lanes per task
  static constexpr int TPW = BT / LPT;               // tasks per workgroup
  static constexpr int EPL = NT * NT / LPT;          // tile elements per lane
  static constexpr int KC = LPT / NT;                // lanes per row in vector products
  static_assert(BT % LPT == 0 && (NT * NT) % LPT == 0 && LPT % NT == 0, "shape");
};

// Tile loads in two steps so that several tiles' loads are in flight at once:
// tfetch issues unconditional loads from clamped addresses into registers
// (nothing when src is NULL); tcommit pins them (after every fetch of the
// group was issued), masks outside n (identity on the diagonal when ident)
// and stores the NT x NT tile to LDS.
template <typename T, int NT, int W>
__device__ __forceinline__ void tfetch(T* v, const T* __restrict__ src, int n, int q) {
  constexpr int LPT = BShape<NT, W>::LPT, EPL = BShape<NT, W>::EPL;
  if (src == nullptr) return;
#pragma unroll
  for (int e = 0; e < EPL; ++e) {
    const int idx = q + e * LPT, r = idx / NT, c = idx % NT;
    v[e] = src[(size_t)min(r, n - 1) * n + min(c, n - 1)];
  }
}

template <typename T, int NT, int W>
__device__ __forceinline__ void tcommit(T* t, T* v, const T* src, int n, int q, bool ident) {
  constexpr int TP = BShape<NT, W>::TP, LPT = BShape<NT, W>::LPT, EPL = BShape<NT, W>::EPL;
  if (src != nullptr) {
#pragma unroll
    for (int e = 0; e < EPL; ++e) pin_reg(v[e]);
  }
#pragma unroll
  for (int e = 0; e < EPL; ++e) {
    const int idx = q + e * LPT, r = idx / NT, c = idx % NT;
    t[r * TP + c] = (src != nullptr && r < n && c < n) ? v[e] : ((ident && r == c) ? T(1) : T(0));
  }
}


// Position (i, j) of the e-th output element tprod hands this thread, in the
// order of its out() calls.
template <typename T, int NT, int W>
__device__ __forceinline__ void tprod_pos(int e, int& i, int& j) {
  if constexpr (NT == 8) {
    const int q = threadIdx.x & 15;
    i = 2 * (q >> 2) + (e >> 1);
    j = 2 * (q & 3) + (e & 1);
  } else {
    constexpr int LPT = BShape<NT, W>::LPT;
    constexpr int TD = NT / 16, WPT = LPT / 64;
    const int l = lane_id();
    const int t = (int)(threadIdx.x % LPT) / 64 + (e >> 2) * WPT;
    i = 16 * (t / TD) + Mfma<T>::row(l, e & 3);
    j = 16 * (t % TD) + Mfma<T>::col(l, e & 3);
  }
}

// Zeros at this thread's tprod output positions: the start of a fill-in
// product (the block receives its only update there).
template <typename T, int NT, int W>
__device__ __forceinline__ void pzero(T* v) {
#pragma unroll
  for (int e = 0; e < BShape<NT, W>::EPL; ++e) v[e] = T(0);
}

// Vector block (n x m, row-major) into a tile of pitch yp; rows >= n zero.
// The loads are unconditional (clamped rows) and pinned before the select:
// a select on each load lets the compiler turn it into a branch with a wait.
template <typename T, int NT, int W>
__device__ __forceinline__ void yload(T* t, int yp, const T* __restrict__ src, int n, int m, int q) {
  constexpr int LPT = BShape<NT, W>::LPT;
  for (int idx = q; idx < NT * m; idx += LPT) {
    const int r = idx / m, c = idx - r * m;
    T v = T(0);
    if (src != nullptr) {
      v = src[(size_t)min(r, n - 1) * m + c];
      pin_reg(v);
    }
    t[r * yp + c] = (r < n) ? v : T(0);
  }
}

// Tile (NT x NT, pitch TP) to an n x n block of HBM, rows and columns below
// n only (no division by the runtime n).
template <typename T, int NT, int W>
__device__ __forceinline__ void tstore(T* __restrict__ dst, const T* t, int n, int q) {
  constexpr int TP = BShape<NT, W>::TP, LPT = BShape<NT, W>::LPT;
#pragma unroll
  for (int idx = q; idx < NT * NT; idx += LPT) {
    const int r = idx / NT, c = idx % NT;
    if (r < n && c < n) dst[(size_t)r * n + c] = t[r * TP + c];
  }
}

// C = A B (NT x NT) with element functors fa(i, k), fb(k, j), plus a second
// product A2 B2 when use2; out(i, j, v) receives every output element this
// thread owns.  Tiles are zero padded to NT, so the inner extent is NT.
// NT = 8: 2x2 register blocks per lane.  NT >= 16: 16x16 MFMA tiles, the
// task's wavefronts taking tiles round robin, inner loop unrolled with two
// accumulators.  For NT >= 16 the predicates guarding a call must be uniform
// per wavefront.
template <typename T, int NT, int W, typename FA, typename FB, typename FA2, typename FB2, typename FO>
__device__ __forceinline__ void tprod(FA fa, FB fb, bool use2, FA2 fa2, FB2 fb2, FO out) {
  if constexpr (NT == 8) {
    const int q = threadIdx.x & 15;
    const int i = 2 * (q >> 2), j = 2 * (q & 3);
    T a00 = T(0), a01 = T(0), a10 = T(0), a11 = T(0);
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      const T x0 = fa(i, k), x1 = fa(i + 1, k), y0 = fb(k, j), y1 = fb(k, j + 1);
      a00 += x0 * y0;
      a01 += x0 * y1;
      a10 += x1 * y0;
      a11 += x1 * y1;
    }
    if (use2) {
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        const T x0 = fa2(i, k), x1 = fa2(i + 1, k), y0 = fb2(k, j), y1 = fb2(k, j + 1);
        a00 += x0 * y0;
        a01 += x0 * y1;
        a10 += x1 * y0;
        a11 += x1 * y1;
      }
    }
    out(i, j, a00);
    out(i, j + 1, a01);
    out(i + 1, j, a10);
    out(i + 1, j + 1, a11);
  } else {
    constexpr int LPT = BShape<NT, W>::LPT;
    constexpr int TD = NT / 16, TT = TD * TD;  // 16x16 tiles per side / per task
    constexpr int WPT = LPT / 64;              // wavefronts per task
    const int l = lane_id();
    const int wt = (int)(threadIdx.x % LPT) / 64;
    const int kq = l >> 4;
#pragma unroll
    for (int tt = 0; tt < TT / WPT; ++tt) {
      const int t = wt + tt * WPT;
      const int i0 = 16 * (t / TD), j0 = 16 * (t % TD);
      const int ii = i0 + (l & 15), jj = j0 + (l & 15);
      typename Mfma<T>::acc_t a0 = acc_zero<T>(), a1 = acc_zero<T>();
#pragma unroll
      for (int k = 0; k < NT; k += 8) {
        a0 = Mfma<T>::mma(fa(ii, k + kq), fb(k + kq, jj), a0);
        a1 = Mfma<T>::mma(fa(ii, k + 4 + kq), fb(k + 4 + kq, jj), a1);
      }
      if (use2) {
#pragma unroll
        for (int k = 0; k < NT; k += 8) {
          a0 = Mfma<T>::mma(fa2(ii, k + kq), fb2(k + kq, jj), a0);
          a1 = Mfma<T>::mma(fa2(ii, k + 4 + kq), fb2(k + 4 + kq, jj), a1);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) out(i0 + Mfma<T>::row(l, r), j0 + Mfma<T>::col(l, r), a0[r] + a1[r]);
    }
  }
}

template <typename T, int NT, int W, typename FA, typename FB, typename FO>
__device__ __forceinline__ void tprod1(FA fa, FB fb, FO out) {
  tprod<T, NT, W>(fa, fb, false, fa, fb, out);
}

// Vector product Y (n x m) = A (n x kend) X (kend x m) plus an optional
// second term.  KC consecutive lanes share row i (i < NT), each summing a
// chunk of NT/KC inner indices, and a butterfly over those lanes completes
// the sum; lane (i, chunk 0) calls out(i, j, v) for i < n, j < m.  Every
// lane of the task must call it.
template <typename T, int NT, int W, typename FA, typename FX, typename FA2, typename FX2, typename FO>
__device__ __forceinline__ void vprod(int n, int m, int kend, FA fa, FX fx, bool use2, int kend2, FA2 fa2, FX2 fx2,
                                      FO out) {
  constexpr int LPT = BShape<NT, W>::LPT, KC = BShape<NT, W>::KC, CS = NT / KC;
  const int q = threadIdx.x % LPT;
  const int i = q / KC, k0 = CS * (q % KC);
  for (int j = 0; j < m; ++j) {
    T acc = T(0);
#pragma unroll
    for (int e = 0; e < CS; ++e)
      if (k0 + e < kend) acc += fa(i, k0 + e) * fx(k0 + e, j);
    if (use2) {
#pragma unroll
      for (int e = 0; e < CS; ++e)
        if (k0 + e < kend2) acc += fa2(i, k0 + e) * fx2(k0 + e, j);
    }
#pragma unroll
    for (int o = KC / 2; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (k0 == 0 && i < n) out(i, j, acc);
  }
}

// Cholesky + inverse of every slot's A tile (lower: L, strict upper: L^-T,
// dinv: 1/L[r][r]); `node` is this thread's task kind (a deferred target task
// has no factor).  (NT = 32 keeps the 16x16-halves factor: the 32-wide pivot
// pairs of tile_chol_inv need registers this kernel does not have.)
template <typename T, int NT, int W>
__device__ __forceinline__ void slots_chol(BSlot<T, NT>* slots, bool node) {
  constexpr int TP = BShape<NT, W>::TP;
  if constexpr (NT == 32) {  // one task per wavefront (node is wavefront-uniform)
    BSlot<T, NT>& S = slots[wave_id()];
    if (node) tile_chol_inv32b<T>(S.A, TP, S.dinv, S.col);
  } else if constexpr (NT == 16) {  // one task per wavefront (node is wavefront-uniform)
    BSlot<T, NT>& S = slots[wave_id()];
    if (node && lane_id() < 32) tile_chol_inv16<T>(S.A, TP, S.dinv, S.col);
  } else {  // four tasks per wavefront
    constexpr int stride = (int)(sizeof(BSlot<T, NT>) / sizeof(T));
    static_assert(sizeof(BSlot<T, NT>) % sizeof(T) == 0, "slot stride");
    BSlot<T, NT>& S = slots[wave_id() * 4];
    tile_chol_inv_grouped<T, 8>(S.A, S.col, S.dinv, stride, TP);
  }
}

template <typename T, int NT, int W>
__device__ __forceinline__ void batched_system(const BatchProblem& P, int op, unsigned char* smem, int yp) {
  constexpr int TP = BShape<NT, W>::TP, LPT = BShape<NT, W>::LPT, TPW = BShape<NT, W>::TPW;
  BSlot<T, NT>* slots = reinterpret_cast<BSlot<T, NT>*>(smem);
  const int slot = threadIdx.x / LPT, q = threadIdx.x % LPT;
  BSlot<T, NT>& S = slots[slot];
  T* Y0 = reinterpret_cast<T*>(slots + TPW) + (size_t)slot * 3 * NT * yp;
  T* Y1 = Y0 + NT * yp;
  T* Y2 = Y1 + NT * yp;
  const int N = P.N, n = P.n, m = P.m;
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  T* L = static_cast<T*>(P.L);
  T* E = static_cast<T*>(P.E);
  T* x = static_cast<T*>(P.x);
  const bool fac = op == SOCU_OP_FACTOR || op == SOCU_OP_FACTOR_AND_SOLVE;
  const bool fwd = op == SOCU_OP_FACTOR_AND_SOLVE || op == SOCU_OP_SOLVE || op == SOCU_OP_FORWARD;
  const bool bwd = op == SOCU_OP_FACTOR_AND_SOLVE || op == SOCU_OP_SOLVE || op == SOCU_OP_BACKWARD;
  const int Lv = recursive_iterations(N);
  T* At = S.A;
  T* Bt = S.B;
  T* Ct = S.C;
  T* Dt = S.D;
  T* dv = S.dinv;
  // Linv[i][k] of a factored tile: its strict upper triangle holds Linv^T
  // (both reads unconditional: a select on a read becomes a branch)
  auto linv = [&](int i, int k) -> T {
    const T u = At[k * TP + i], d = dv[i];
    return (k < i) ? u : ((k == i) ? d : T(0));
  };

  if (fac || fwd) {
    int offp = 0, off = 0;  // E segment offsets of levels l-1 and l
    for (int l = 0; l < Lv; ++l) {
      const int s = 1 << l, h = s >> 1;
      if (l > 0) {
        offp = off;
        off = off + (N >> (l - 1)) - 1;
      }
      const int dim = (N + s) / (2 * s);
      const int nB = (l > 0) ? N / (2 * s) : 0;
      const int ntask = dim + nB;
      for (int base = 0; base < ntask; base += TPW) {
        const int t = base + slot;
        const bool node = t < dim;
        const bool defer = !node && t < ntask;
        const bool task = node || defer;
        const int k = node ? s - 1 + 2 * s * t : 2 * s * (t - dim + 1) - 1;  // block row written
        const bool up = l > 0 && task;       // receives from level l-1
        const bool right = up && k + h < N;  // Lc_{k+h} exists
        const bool has_lc = node && k >= s;  // Lc_k exists (and, for l > 0, Lc_{k-h})
        const bool has_r = node && k + s < N;  // R_k exists (and, for l > 0, R_{k+h})
        const int sh = l - 1;                  // log2(h)
        T* Lk = task ? L + (size_t)k * nn : nullptr;
        T* Rk = has_r ? E + (size_t)(off + k / s) * nn : nullptr;
        T* Lck = has_lc ? E + (size_t)(off + (k - s) / s) * nn : nullptr;
        const T* Rl = up ? E + (size_t)(offp + ((k - h) >> sh)) * nn : nullptr;                // (k, k-h)
        const T* Lcr = right ? E + (size_t)(offp + (k >> sh)) * nn : nullptr;                  // (k+h, k)
        const T* Lcl = (up && has_lc) ? E + (size_t)(offp + ((k - s) >> sh)) * nn : nullptr;   // (k-h, k-s)
        const T* Rr = (up && has_r) ? E + (size_t)(offp + ((k + h) >> sh)) * nn : nullptr;     // (k+s, k+h)
        T* yk = (fwd && task) ? x + (size_t)k * nm : nullptr;

        // ---- P1: loads, every tile's (and the vectors') in flight at once ----
        constexpr int EPL = BShape<NT, W>::EPL;  // tile / product elements per lane
        {
          // level 0 (node): B <- R_k, D <- Lc_k (original couplings);
          // level > 0: B <- R_{k-h}, C <- Lc_{k+h}, D <- Lc_{k-h}
          const T* srcB = up ? Rl : ((node && fac) ? Rk : nullptr);
          const T* srcD = up ? (fac ? Lcl : nullptr) : ((node && fac) ? Lck : nullptr);
          auto ydv = [&]() {  // vectors, and 1/L[q][q] for solves of a stored factor
            if (fwd) {
              yload<T, NT, W>(Y0, yp, yk, n, m, q);
              if (up) {
                yload<T, NT, W>(Y1, yp, x + (size_t)(k - h) * nm, n, m, q);
                yload<T, NT, W>(Y2, yp, right ? x + (size_t)(k + h) * nm : nullptr, n, m, q);
              }
            }
            if (!fac && node && q < NT) {
              T d = Lk[(size_t)min(q, n - 1) * (n + 1)];
              pin_reg(d);
              dv[q] = (q < n) ? T(1) / d : T(1);
            }
          };
          const T* srcC = up ? Lcr : nullptr;
          if (!(fac && (up || node))) srcD = nullptr;
          if constexpr (EPL * sizeof(T) <= 64) {  // every load in flight at once
            T va[EPL], vb[EPL], vc[EPL], vd[EPL];
            tfetch<T, NT, W>(va, Lk, n, q);
            tfetch<T, NT, W>(vb, srcB, n, q);
            tfetch<T, NT, W>(vc, srcC, n, q);
            tfetch<T, NT, W>(vd, srcD, n, q);
            ydv();
            tcommit<T, NT, W>(At, va, Lk, n, q, true);
            if (up || (node && fac)) tcommit<T, NT, W>(Bt, vb, srcB, n, q, false);
            if (up) tcommit<T, NT, W>(Ct, vc, srcC, n, q, false);
            if (fac && (up || node)) tcommit<T, NT, W>(Dt, vd, srcD, n, q, false);
          } else {
            // FP64 tiles of 16 elements per lane (NT = 32, one wavefront): one
            // tile at a time; two in flight spill registers in this kernel
            auto tl = [&](T* t, const T* sp, bool ident) {
              T v[EPL];
              tfetch<T, NT, W>(v, sp, n, q);
              tcommit<T, NT, W>(t, v, sp, n, q, ident);
            };
            tl(At, Lk, true);
            if (up || (node && fac)) tl(Bt, srcB, false);
            if (up) tl(Ct, srcC, false);
            if (fac && (up || node)) tl(Dt, srcD, false);
            ydv();
          }
        }
        // level > 0: node k's couplings are fill-in blocks that receive their
        // only update here (the reference's F -= R Lc on a zeroed block), so
        // the fill products start from zero and the fill-in region's contents
        // on entry are never read
        T old[EPL];
        __syncthreads();

        // ---- P2: updates from level l-1 ----
        int nk = 0;
        if (up) {
          if (fac)
            tprod<T, NT, W>(
                [&](int i, int kk) { return Bt[i * TP + kk]; }, [&](int kk, int j) { return Bt[j * TP + kk]; },
                right, [&](int i, int kk) { return Ct[kk * TP + i]; },
                [&](int kk, int j) { return Ct[kk * TP + j]; }, [&](int i, int j, T v) { At[i * TP + j] -= v; });
          if (fwd)
            vprod<T, NT, W>(
                n, m, n, [&](int i, int kk) { return Bt[i * TP + kk]; }, [&](int kk, int j) { return Y1[kk * yp + j]; },
                right, n, [&](int i, int kk) { return Ct[kk * TP + i]; },
                [&](int kk, int j) { return Y2[kk * yp + j]; }, [&](int i, int j, T v) { Y0[i * yp + j] -= v; });
          if (fac && has_lc) {
            pzero<T, NT, W>(old);
            tprod1<T, NT, W>(
                [&](int i, int kk) { return Bt[i * TP + kk]; }, [&](int kk, int j) { return Dt[kk * TP + j]; },
                [&](int i, int j, T v) {
                  old[nk] = (i < n && j < n) ? old[nk] - v : T(0);
                  ++nk;
                });
          }
        }
        __syncthreads();

        // ---- P3: finish the fill-in of Lc_k, stage R_{k+h}; deferred stores ----
        if (up && fac && has_lc)
#pragma unroll
          for (int e = 0; e < EPL; ++e) {
            int i, j;
            tprod_pos<T, NT, W>(e, i, j);
            Dt[i * TP + j] = old[e];
          }
        if (up && fac) {  // stage R_{k+h}; fetch the old R_k with it
          T vb[EPL];
          tfetch<T, NT, W>(vb, Rr, n, q);
          tcommit<T, NT, W>(Bt, vb, Rr, n, q, false);
          if (has_r) pzero<T, NT, W>(old);
        }
        if (defer) {
          if (fac) tstore<T, NT, W>(Lk, At, n, q);
          if (fwd)
            for (int idx = q; idx < n * m; idx += LPT) yk[idx] = Y0[(idx / m) * yp + idx % m];
        }
        __syncthreads();

        // ---- P4: fill-in of R_k (workgroup-uniform condition: barriers) ----
        if (l > 0 && fac) {
          nk = 0;
          if (has_r)
            tprod1<T, NT, W>(
                [&](int i, int kk) { return Bt[i * TP + kk]; }, [&](int kk, int j) { return Ct[kk * TP + j]; },
                [&](int i, int j, T v) {
                  old[nk] = (i < n && j < n) ? old[nk] - v : T(0);
                  ++nk;
                });
          __syncthreads();
          if (has_r)
#pragma unroll
            for (int e = 0; e < EPL; ++e) {
              int i, j;
              tprod_pos<T, NT, W>(e, i, j);
              Bt[i * TP + j] = old[e];
            }
          __syncthreads();
        }

        // ---- P5: factor and invert the diagonal blocks ----
        if (fac) {
          slots_chol<T, NT, W>(slots, node);
          __syncthreads();
        }

        // ---- P6: node outputs ----
        if (node) {
          if (fac) {
            tstore<T, NT, W>(Lk, At, n, q);
            if (has_lc)  // Lc_k <- Linv_k Lc_k
              tprod1<T, NT, W>(
                  linv, [&](int kk, int j) { return Dt[kk * TP + j]; },
                  [&](int i, int j, T v) {
                    if (i < n && j < n) Lck[(size_t)i * n + j] = v;
                  });
            if (has_r)  // R_k <- R_k Linv_k^T
              tprod1<T, NT, W>(
                  [&](int i, int kk) { return Bt[i * TP + kk]; }, [&](int kk, int j) { return linv(j, kk); },
                  [&](int i, int j, T v) {
                    if (i < n && j < n) Rk[(size_t)i * n + j] = v;
                  });
          }
          if (fwd)  // y_k <- Linv_k y_k
            vprod<T, NT, W>(
                n, m, n, linv, [&](int kk, int j) { return Y0[kk * yp + j]; }, false, 0, linv,
                [&](int kk, int j) { return Y0[kk * yp + j]; }, [&](int i, int j, T v) { yk[(size_t)i * m + j] = v; });
        }
        __syncthreads();
      }
    }
  }

  if (bwd) {
    for (int l = Lv - 1; l >= 0; --l) {
      const int s = 1 << l;
      const int dim = (N + s) / (2 * s);
      int off = 0;
      for (int u = 0; u < l; ++u) off += (N >> u) - 1;
      for (int base = 0; base < dim; base += TPW) {
        const int t = base + slot;
        const bool node = t < dim;
        const int i = s - 1 + 2 * s * t;
        const bool right = node && i + s < N, left = node && i >= s;
        T* Li = node ? L + (size_t)i * nn : nullptr;
        T* xi = node ? x + (size_t)i * nm : nullptr;
        {
          const T* Ri = right ? E + (size_t)(off + i / s) * nn : nullptr;       // R_i
          const T* Lci = left ? E + (size_t)(off + (i - s) / s) * nn : nullptr;  // Lc_i
          constexpr bool kAll = BShape<NT, W>::EPL * sizeof(T) <= 64;
          T va[BShape<NT, W>::EPL], vb[BShape<NT, W>::EPL], vc[BShape<NT, W>::EPL];
          tfetch<T, NT, W>(va, Li, n, q);
          tfetch<T, NT, W>(vb, Ri, n, q);
          if constexpr (!kAll) {
            tcommit<T, NT, W>(At, va, Li, n, q, true);
            tcommit<T, NT, W>(Bt, vb, Ri, n, q, false);
            asm volatile("" ::: "memory");
          }
          tfetch<T, NT, W>(vc, Lci, n, q);
          yload<T, NT, W>(Y0, yp, xi, n, m, q);
          yload<T, NT, W>(Y1, yp, right ? x + (size_t)(i + s) * nm : nullptr, n, m, q);
          yload<T, NT, W>(Y2, yp, left ? x + (size_t)(i - s) * nm : nullptr, n, m, q);
          if (node && q < NT) {
            T d = Li[(size_t)min(q, n - 1) * (n + 1)];
            pin_reg(d);
            dv[q] = (q < n) ? T(1) / d : T(1);
          }
          if constexpr (kAll) {
            tcommit<T, NT, W>(At, va, Li, n, q, true);
            tcommit<T, NT, W>(Bt, vb, Ri, n, q, false);
          }
          tcommit<T, NT, W>(Ct, vc, Lci, n, q, false);
        }
        __syncthreads();
        // x_i - R_i^T x_{i+s} - Lc_i x_{i-s} (in place: element (r, j) of Y0 is
        // read and written by one lane only)
        if (node)
          vprod<T, NT, W>(
              n, m, n, [&](int r, int kk) { return Bt[kk * TP + r]; }, [&](int kk, int j) { return Y1[kk * yp + j]; },
              true, n, [&](int r, int kk) { return Ct[r * TP + kk]; }, [&](int kk, int j) { return Y2[kk * yp + j]; },
              [&](int r, int j, T v) { Y0[r * yp + j] -= v; });
        __syncthreads();
        // x_i <- Linv_i^T x_i
        if (node)
          vprod<T, NT, W>(
              n, m, n, [&](int r, int kk) { return linv(kk, r); }, [&](int kk, int j) { return Y0[kk * yp + j]; },
              false, 0, [&](int r, int kk) { return linv(kk, r); }, [&](int kk, int j) { return Y0[kk * yp + j]; },
              [&](int r, int j, T v) { xi[(size_t)r * m + j] = v; });
        __syncthreads();
      }
    }
  }
}

// One workgroup per system of tile class NT; `order` lists the systems.
template <typename T, int NT, int W>
__global__ __launch_bounds__(64 * W) void k_batched(const BatchProblem* __restrict__ probs,
                                                    const int* __restrict__ order, const uint8_t* __restrict__ mask,
                                                    int op, int yp) {
  const int p = order[blockIdx.x];
  if (masked_out(mask, p)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const BatchProblem P = probs[p];
  batched_system<T, NT, W>(P, op, smem_raw, yp);
}

template <typename T, int NT, int W>
size_t lds_bytes(int yp) {
  return BShape<NT, W>::TPW * (sizeof(BSlot<T, NT>) + (size_t)3 * NT * yp * sizeof(T));
}

constexpr size_t kLdsPerCu = 160 * 1024;

// (NT = 32, FP64, four wavefronts: four tasks' tiles leave room for m <= 7)
template <typename T, int NT, int W>
hipError_t set_attr() {
  const size_t lds = lds_bytes<T, NT, W>(NT + 1);
  return hipFuncSetAttribute((const void*)k_batched<T, NT, W>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(lds < kLdsPerCu ? lds : kLdsPerCu));
}

template <typename T, int NT>
bool fits_nt(int waves, int mmax) {
  const int yp = mmax + 1;
  const size_t lds = waves == 1 ? lds_bytes<T, NT, 1>(yp) : (waves == 2 ? lds_bytes<T, NT, 2>(yp) : lds_bytes<T, NT, 4>(yp));
  return lds <= kLdsPerCu;
}

template <typename T, int NT>
hipError_t set_attrs() {
  hipError_t e = set_attr<T, NT, 1>();
  if (e == hipSuccess) e = set_attr<T, NT, 2>();
  return e == hipSuccess ? set_attr<T, NT, 4>() : e;
}

template <typename T, int NT>
void launch_nt(const BatchProblem* dprobs, const int* dorder, int count, int mmax, int waves, const uint8_t* mask,
               int op, hipStream_t st) {
  const int yp = mmax + 1;
  if (waves == 1) {
    const size_t lds = lds_bytes<T, NT, 1>(yp);
    hipLaunchKernelGGL((k_batched<T, NT, 1>), dim3(count), dim3(64), lds, st, dprobs, dorder, mask, op, yp);
  } else if (waves == 2) {
    const size_t lds = lds_bytes<T, NT, 2>(yp);
    hipLaunchKernelGGL((k_batched<T, NT, 2>), dim3(count), dim3(128), lds, st, dprobs, dorder, mask, op, yp);
  } else {
    const size_t lds = lds_bytes<T, NT, 4>(yp);
    hipLaunchKernelGGL((k_batched<T, NT, 4>), dim3(count), dim3(256), lds, st, dprobs, dorder, mask, op, yp);
  }
}

// Copies every system's L, E and x from src to dst (same shapes); one
// workgroup per system.
template <typename T>
__global__ __launch_bounds__(256) void k_copy_systems(const BatchProblem* __restrict__ src,
                                                      const BatchProblem* __restrict__ dst, int op) {
  const BatchProblem a = src[blockIdx.x], b = dst[blockIdx.x];
  const size_t nn = (size_t)a.n * a.n;
  const size_t nL = (size_t)a.N * nn, nE = (size_t)off_diag_storage_len(a.N) * nn;
  const size_t nx = (op == SOCU_OP_FACTOR) ? 0 : (size_t)a.N * a.n * a.m;
  for (size_t k = threadIdx.x; k < nL; k += blockDim.x) static_cast<T*>(b.L)[k] = static_cast<const T*>(a.L)[k];
  for (size_t k = threadIdx.x; k < nE; k += blockDim.x) static_cast<T*>(b.E)[k] = static_cast<const T*>(a.E)[k];
  for (size_t k = threadIdx.x; k < nx; k += blockDim.x) static_cast<T*>(b.x)[k] = static_cast<const T*>(a.x)[k];
}

}  // namespace

hipError_t init_batched_kernels() {
  hipError_t e = set_attrs<double, 8>();
  if (e == hipSuccess) e = set_attrs<double, 16>();
  if (e == hipSuccess) e = set_attrs<double, 32>();
  if (e == hipSuccess) e = set_attrs<float, 8>();
  if (e == hipSuccess) e = set_attrs<float, 16>();
  if (e == hipSuccess) e = set_attrs<float, 32>();
  return e;
}

int batched_class(int n) { return n <= 8 ? 0 : (n <= 16 ? 1 : 2); }

bool batched_fits(int cls, int waves, int mmax, int dtype) {
  if (cls < 0 || cls > 2 || (waves != 1 && waves != 2 && waves != 4)) return false;
  if (dtype == SOCU_DTYPE_F64)
    return cls == 0 ? fits_nt<double, 8>(waves, mmax) : (cls == 1 ? fits_nt<double, 16>(waves, mmax) : fits_nt<double, 32>(waves, mmax));
  return cls == 0 ? fits_nt<float, 8>(waves, mmax) : (cls == 1 ? fits_nt<float, 16>(waves, mmax) : fits_nt<float, 32>(waves, mmax));
}

hipError_t batched_launch_class(const BatchProblem* dprobs, const int* dorder, int cls, int count, int mmax, int waves,
                                const uint8_t* mask, int op, int dtype, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (cls < 0 || cls > 2 || mmax < 1 || mmax > (8 << cls) || !batched_fits(cls, waves, mmax, dtype))
    return hipErrorInvalidValue;
  if (dtype == SOCU_DTYPE_F64) {
    if (cls == 0) launch_nt<double, 8>(dprobs, dorder, count, mmax, waves, mask, op, st);
    else if (cls == 1) launch_nt<double, 16>(dprobs, dorder, count, mmax, waves, mask, op, st);
    else launch_nt<double, 32>(dprobs, dorder, count, mmax, waves, mask, op, st);
  } else {
    if (cls == 0) launch_nt<float, 8>(dprobs, dorder, count, mmax, waves, mask, op, st);
    else if (cls == 1) launch_nt<float, 16>(dprobs, dorder, count, mmax, waves, mask, op, st);
    else launch_nt<float, 32>(dprobs, dorder, count, mmax, waves, mask, op, st);
  }
  return hipGetLastError();
}

hipError_t batched_copy(const BatchProblem* dsrc, const BatchProblem* ddst, int count, int op, int dtype,
                        hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (dtype == SOCU_DTYPE_F64)
    hipLaunchKernelGGL(k_copy_systems<double>, dim3(count), dim3(256), 0, st, dsrc, ddst, op);
  else
    hipLaunchKernelGGL(k_copy_systems<float>, dim3(count), dim3(256), 0, st, dsrc, ddst, op);
  return hipGetLastError();
}

}  // namespace socu
