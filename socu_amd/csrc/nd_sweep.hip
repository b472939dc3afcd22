// Persistent dataflow solve sweeps for one right-hand side (m = 1), n <= 128.
//
// The level-synchronous sweeps (k_sweep32 / k_sweep_nb / k_apply_fwd)
// need a launch per level and direction, each waiting for the slowest
// workgroup of its level and paying a dispatch gap; at the upper levels, where
// a level is one or two nodes, those gaps are the whole cost.  Here one launch
// runs both sweeps as node tasks that wait on per-node counters:
//
//   forward task of node j (level l, s = 2^l), reference forward sweep
//   block_tridiag_solver.py:425-485 (oracle nd.forward):
//       y_j <- L_j^{-1} (b_j - sum of the contributions received)
//       contribution R_j y_j      -> node j+s   (R_j = E[off_l + j/s])
//       contribution Lc_j^T y_j   -> node j-s   (Lc_j = E[off_l + (j-s)/s])
//   backward task of node i (block_tridiag_solver.py:581-643, nd.backward):
//       x_i <- L_i^{-T} (y_i - R_i^T x_{i+s} - Lc_i x_{i-s})
//
// Contributions are not accumulated with atomics: each goes to its own slot
// (step l' of target j, side 0 from j-s', side 1 from j+s'), and the target
// sums its slots in a fixed order, so results are reproducible run to run.
//
// Scheduling: workgroups take tasks from a ticket counter in the order
// forward level 0, 1, ..., top, then backward top, ..., 0.  A task only waits
// on tasks with smaller tickets, which were handed to workgroups that are
// already running, so the sweep cannot deadlock whatever the grid size or
// residency.  Each task loads its three blocks (L_j with the stored inverse,
// R_j, Lc_j) into registers before it waits, so only the n-vectors are on the
// dependency chain.  Every wait is bounded: on expiry the kernel records a
// status word, stops waiting and drains (results are then undefined).
//
// Layouts of a block in registers (NP = 32*NB, THREADS = NP*H):
//   column layout: thread (c = t % NP, h = t / NP) holds rows h + H*e,
//                  column c; column products out[c] = sum_r A[r][c] v[r]
//                  reduce H partials through LDS.
//   row layout:    8 lanes per row, KPL = NP/8 consecutive columns per lane;
//                  row products out[r] = sum_k A[r][k] v[k] reduce with three
//                  lane shuffles.
#include <stdlib.h>
#include <string.h>

#include <atomic>

#include "nd_common.h"
#include "nd_internal.h"
#include "nd_tile.h"

namespace socu {

namespace {

template <typename T, int NB>
struct SweepShape {
  static constexpr int NP = 32 * NB;
  // column layout row groups: 8 for n <= 32 and for FP64 n > 96 (1024
  // threads, so that two FP64 blocks fit the registers), else 4
  static constexpr int H = (NB == 1 || (sizeof(T) == 8 && NB == 4)) ? 8 : 4;
  static constexpr int THREADS = NP * H;  // 256, 256, 384, 512 (FP64: 1024)
  static constexpr int EPT = NP * NP / THREADS;
  static constexpr int G = 8;                // row layout lanes per row
  static constexpr int KPL = NP / G;
  static constexpr int RP = THREADS / G;     // rows per pass
  static constexpr int PASSES = NP / RP;
  static_assert(EPT == KPL * PASSES, "layouts hold the same number of elements");
};

// A zero the compiler cannot see through.  Index and mask arithmetic below
// depends on it, so it is evaluated where it is used: otherwise the per-element
// offsets of both layouts are hoisted out of the task loop (and the mask
// compares above the loads), each holding a register for the whole kernel.
__device__ __forceinline__ int opaque_zero() {
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// A[idx] through a 32-bit byte offset from a uniform base (idx < 2^16 for
// n <= 128), which the loads take as scalar base + vector offset instead of a
// 64-bit address per element.
template <typename T>
__device__ __forceinline__ T ld_off(const T* A, int idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(A) + (unsigned)(idx * (int)sizeof(T)));
}

// Column layout: v[e] = A[h + H e][c] (clamped address; masked later).
template <typename T, int NB>
__device__ __forceinline__ void load_cl(T* v, const T* __restrict__ A, int n, int t) {
  using S = SweepShape<T, NB>;
  const int z = opaque_zero();
  const int c = t % S::NP + z, h = t / S::NP + z;
#pragma unroll
  for (int e = 0; e < S::EPT; ++e) {
    const int r = h + S::H * e;
    v[e] = ld_off(A, min(r, n - 1) * n + min(c, n - 1));
  }
}

// Row layout: v[p*KPL + e] = A[rr + RP p][sub*KPL + e].
template <typename T, int NB>
__device__ __forceinline__ void load_rl(T* v, const T* __restrict__ A, int n, int t) {
  using S = SweepShape<T, NB>;
  const int z = opaque_zero();
  const int sub = t % S::G + z, rr = t / S::G + z;
#pragma unroll
  for (int p = 0; p < S::PASSES; ++p)
#pragma unroll
    for (int e = 0; e < S::KPL; ++e) {
      const int r = rr + S::RP * p, k = sub * S::KPL + e;
      v[p * S::KPL + e] = ld_off(A, min(r, n - 1) * n + min(k, n - 1));
    }
}

// Zero the column-layout entries outside the n x n block (and, with
// `strict_upper`, those with r >= c).
template <typename T, int NB>
__device__ __forceinline__ void mask_cl(T* v, int n, int t, bool strict_upper) {
  using S = SweepShape<T, NB>;
  const int z = opaque_zero();
  const int c = t % S::NP + z, h = t / S::NP + z;
#pragma unroll
  for (int e = 0; e < S::EPT; ++e) {
    const int r = h + S::H * e;
    const bool ok = r < n && c < n && (!strict_upper || r < c);
    v[e] = ok ? v[e] : T(0);
  }
}

template <typename T, int NB>
__device__ __forceinline__ void mask_rl(T* v, int n, int t, bool strict_upper) {
  using S = SweepShape<T, NB>;
  const int z = opaque_zero();
  const int sub = t % S::G + z, rr = t / S::G + z;
#pragma unroll
  for (int p = 0; p < S::PASSES; ++p)
#pragma unroll
    for (int e = 0; e < S::KPL; ++e) {
      const int r = rr + S::RP * p, k = sub * S::KPL + e;
      const bool ok = r < n && k < n && (!strict_upper || k > r);
      v[p * S::KPL + e] = ok ? v[p * S::KPL + e] : T(0);
    }
}

// part[h*NP + c] = sum_e v[e] y[h + H e]   (column layout, y in LDS)
template <typename T, int NB>
__device__ __forceinline__ void prod_cl(const T* v, const T* y, T* part, int t) {
  using S = SweepShape<T, NB>;
  const int z = opaque_zero();
  const int c = t % S::NP + z, h = t / S::NP + z;
  T a0 = T(0), a1 = T(0);
#pragma unroll
  for (int e = 0; e < S::EPT; e += 2) {
    a0 += v[e] * y[h + S::H * e];
    a1 += v[e + 1] * y[h + S::H * (e + 1)];
  }
  part[h * S::NP + c] = a0 + a1;
}

// out(r, sum_k A[r][k] y[k]) on the lanes with sub == 0   (row layout)
template <typename T, int NB, typename F>
__device__ __forceinline__ void prod_rl(const T* v, const T* y, int t, F out) {
  using S = SweepShape<T, NB>;
  const int z = opaque_zero();
  const int sub = t % S::G + z, rr = t / S::G + z;
#pragma unroll
  for (int p = 0; p < S::PASSES; ++p) {
    T a0 = T(0), a1 = T(0);
#pragma unroll
    for (int e = 0; e < S::KPL; e += 2) {
      a0 += v[p * S::KPL + e] * y[sub * S::KPL + e];
      a1 += v[p * S::KPL + e + 1] * y[sub * S::KPL + e + 1];
    }
    T acc = a0 + a1;
#pragma unroll
    for (int o = S::G / 2; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, S::G);
    if (sub == 0) out(rr + S::RP * p, acc);
  }
}

// Nodes of level l: i = s-1+2s*t for t < (N+s)/(2s), s = 2^l.
__device__ __forceinline__ int sweep_level_dim(int N, int l) { return (N + (1 << l)) >> (l + 1); }

// Ticket -> (level, direction, index within the level's tasks).  Phases:
// forward levels lmin..Lv-1 (if fwd), then backward levels Lv-1..lmin (if
// bwd); levels below lmin are swept by the level kernels around this launch.
__device__ __forceinline__ bool sweep_decode(int task, int B, int N, int Lv, int lmin, bool fwd, bool bwd, int& l,
                                             bool& back, int& q) {
  if (fwd)
    for (int k = lmin; k < Lv; ++k) {
      const int cnt = sweep_level_dim(N, k) * B;
      if (task < cnt) {
        l = k;
        back = false;
        q = task;
        return true;
      }
      task -= cnt;
    }
  if (bwd)
    for (int k = Lv - 1; k >= lmin; --k) {
      const int cnt = sweep_level_dim(N, k) * B;
      if (task < cnt) {
        l = k;
        back = true;
        q = task;
        return true;
      }
      task -= cnt;
    }
  return false;
}

// Values passed between tasks (contribution slots, y_j, x_j) are written and
// read with ld_coh / st_coh (nd_common.h): a producer waits for its stores to
// complete before raising a counter, a consumer reads after seeing it, with
// no release fence (which writes back the whole L2: 1 us for a lone task,
// 7-9 us per task with 32-64 tasks in flight) and no acquire invalidation.

// Bounded wait (one thread): until *p >= target, at most `limit` polls.
// Returns false, and raises the status word (and the host-visible word
// hstatus, if any), when the bound expires or an earlier wait already
// expired.  limit < 0 makes every wait that is not already satisfied expire
// at once (a host-side switch for tests of this path; every kernel still
// exits normally).
__device__ bool wait_at_least(const int* p, int target, int* status, int* hstatus, int limit) {
  int spins = 0;
  while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    if (limit < 0 || ++spins > limit) {
      __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (hstatus != nullptr) __hip_atomic_store(hstatus, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// U block in LDS: strict upper triangle of L_j (the stored inverse,
// U[r][c] = Linv[c][r] for r < c), zero elsewhere; pitch NP + 1.
template <typename T, int NB>
__device__ __forceinline__ void store_upper(T* Us, const T* v, int n, int t) {
  using S = SweepShape<T, NB>;
  const int z = opaque_zero();
  const int c = t % S::NP + z, h = t / S::NP + z;
#pragma unroll
  for (int e = 0; e < S::EPT; ++e) {
    const int r = h + S::H * e;
    Us[r * (S::NP + 1) + c] = (r < c && c < n) ? v[e] : T(0);
  }
}

// part[h*NP + c] = sum_e Us[h + H e][c] y[h + H e]      (U^T y partials)
template <typename T, int NB>
__device__ __forceinline__ void prod_upper_t(const T* Us, const T* y, T* part, int t) {
  using S = SweepShape<T, NB>;
  const int z = opaque_zero();
  const int c = t % S::NP + z, h = t / S::NP + z;
  T a0 = T(0), a1 = T(0);
#pragma unroll
  for (int e = 0; e < S::EPT; e += 2) {
    a0 += Us[(h + S::H * e) * (S::NP + 1) + c] * y[h + S::H * e];
    a1 += Us[(h + S::H * (e + 1)) * (S::NP + 1) + c] * y[h + S::H * (e + 1)];
  }
  part[h * S::NP + c] = a0 + a1;
}

// part[h*NP + r] = sum_e Us[r][h + H e] y[h + H e]      (U y partials)
template <typename T, int NB>
__device__ __forceinline__ void prod_upper(const T* Us, const T* y, T* part, int t) {
  using S = SweepShape<T, NB>;
  const int z = opaque_zero();
  const int r = t % S::NP + z, h = t / S::NP + z;
  T a0 = T(0), a1 = T(0);
#pragma unroll
  for (int e = 0; e < S::EPT; e += 2) {
    a0 += Us[r * (S::NP + 1) + h + S::H * e] * y[h + S::H * e];
    a1 += Us[r * (S::NP + 1) + h + S::H * (e + 1)] * y[h + S::H * (e + 1)];
  }
  part[h * S::NP + r] = a0 + a1;
}

template <typename T, int NB>
constexpr size_t sweep_lds_bytes() {
  return sizeof(T) * (size_t)(32 * NB) * (32 * NB + 1);
}

template <typename T, int NB>
__global__ __launch_bounds__((SweepShape<T, NB>::THREADS)) void k_sweep(const T* __restrict__ L, const T* __restrict__ E,
                                                                   T* x, const uint8_t* __restrict__ mask, int* ws,
                                                                   T* slots, int* hstatus, int spin_limit, int B,
                                                                   int N, int n, int dirs, int lmin) {
  using S = SweepShape<T, NB>;
  constexpr int NP = S::NP;
  // Three blocks in flight at once when they fit in registers; otherwise the
  // L block is loaded and parked in LDS before R and Lc are loaded.
  constexpr bool kOverlap = sizeof(T) * NB <= 16;
  const bool fwd = (dirs & 1) != 0, bwd_on = (dirs & 2) != 0;
  const int Lv = recursive_iterations(N);
  const int nE = (int)off_diag_storage_len(N);
  const int nslot = sweep_slots_per_system(N);
  const size_t nn = (size_t)n * n;
  int* ticket = ws;
  int* status = ws + 1;
  int* fcount = ws + kSweepHeader;
  int* fdone = fcount + (size_t)B * N;
  int* bdone = fdone + (size_t)B * N;

  extern __shared__ __align__(16) unsigned char sweep_smem[];
  T* Us = reinterpret_cast<T*>(sweep_smem);
  __shared__ T Y[NP], Y2[NP], XA[NP], XC[NP], D[NP], RW[NP];
  __shared__ T part[S::H * NP], part2[S::H * NP];
  __shared__ int s_task, s_ok;
  const int t = threadIdx.x;

  // With every system masked out (a refinement step after convergence) the
  // launch ends at its first barrier; the mask is read while the first ticket
  // is taken.  (Masked tasks touch no state, so neither does this exit.)
  bool first = true;
  for (;;) {
    if (t == 0) s_task = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int any = 1;
    if (first && mask != nullptr) {
      any = 0;
      for (int b = t; b < B; b += S::THREADS) any |= masked_out(mask, b) ? 0 : 1;
    }
    first = false;
    if (!__syncthreads_or(any)) break;
    int l = 0, q = 0;
    bool bwd = false;
    const int task = __builtin_amdgcn_readfirstlane(s_task);  // uniform: scalar decode and addressing
    if (!sweep_decode(task, B, N, Lv, lmin, fwd, bwd_on, l, bwd, q)) break;
    const int tt = q / B, b = q - tt * B;
    const int s = 1 << l;
    const int j = s - 1 + 2 * s * tt;
    // E segment of level l and first contribution slot of step l
    int offl = 0, sofl = 0;
    for (int k = 0; k < l; ++k) {
      offl += (N >> k) - 1;
      sofl += 2 * (N >> (k + 1));
    }
    if (!masked_out(mask, b)) {
      const T* U = L + ((size_t)b * N + j) * nn;
      const bool hasR = j + s < N, hasC = j >= s;
      const T* Rb = hasR ? E + ((size_t)b * nE + offl + j / s) * nn : U;
      const T* Cb = hasC ? E + ((size_t)b * nE + offl + (j - s) / s) * nn : U;
      T* xj = x + ((size_t)b * N + j) * n;
      T u[S::EPT], r[S::EPT], c[S::EPT];
      // Every block load is issued before the dependency wait; the pins keep
      // them together and ahead of the wait.
      const T dg = ld_off(U, min(t, n - 1) * n + min(t, n - 1));
      load_cl<T, NB>(u, U, n, t);
      if (!kOverlap) {
#pragma unroll
        for (int e = 0; e < S::EPT; ++e) pin_reg(u[e]);
        store_upper<T, NB>(Us, u, n, t);
        asm volatile("" ::: "memory");  // R and Lc loads stay below: u's registers are free by then
      }
      if (!bwd) {
        load_rl<T, NB>(r, Rb, n, t);
        load_cl<T, NB>(c, Cb, n, t);
      } else {
        load_cl<T, NB>(r, Rb, n, t);
        load_rl<T, NB>(c, Cb, n, t);
      }
#pragma unroll
      for (int e = 0; e < S::EPT; ++e) {
        pin_reg(r[e]);
        pin_reg(c[e]);
        if (kOverlap) pin_reg(u[e]);
      }
      pin_reg(const_cast<T&>(dg));
      if (kOverlap) store_upper<T, NB>(Us, u, n, t);
      if (t < NP) D[t] = t < n ? dg : T(1);
      if (!bwd) {
        mask_rl<T, NB>(r, n, t, false);
        mask_cl<T, NB>(c, n, t, false);
      } else {
        mask_cl<T, NB>(r, n, t, false);
        mask_rl<T, NB>(c, n, t, false);
      }

      if (t == 0) {
        bool ok = true;
        if (!bwd) {
          int need = 0;
          for (int lp = lmin; lp < l; ++lp) need += 1 + ((j + (1 << lp) < N) ? 1 : 0);
          if (need > 0) ok = wait_at_least(fcount + (size_t)b * N + j, need, status, hstatus, spin_limit);
        } else {
          if (fwd) ok = wait_at_least(fdone + (size_t)b * N + j, 1, status, hstatus, spin_limit);
          if (ok && hasR) ok = wait_at_least(bdone + (size_t)b * N + j + s, 1, status, hstatus, spin_limit);
          if (ok && hasC) ok = wait_at_least(bdone + (size_t)b * N + j - s, 1, status, hstatus, spin_limit);
        }
        s_ok = ok ? 1 : 0;
      }
      __syncthreads();
      // After an expired wait this task's inputs may be incomplete: its
      // output is NaN (and so is everything computed from it), so the failure
      // shows in the solution, besides the status words.
      const bool ok = s_ok != 0;

      if (!bwd) {
        // y_j = b_j - contributions, summed in a fixed order
        // (contributions of levels below lmin are already in x_j)
        {
          // four thread groups sum the slots of levels lp = lmin + g (mod 4)
          const int g = t / NP, tc = t % NP;
          T v = T(0);
          if (g < 4 && tc < n) {
            const T* sb = slots + (size_t)b * nslot * n;
            int so = 0;
            for (int k = 0; k < lmin + g && k < l; ++k) so += 2 * (N >> (k + 1));
#pragma unroll 2
            for (int lp = lmin + g; lp < l; lp += 4) {
              const int sp = 1 << lp;
              const int u0 = (j + 1) / (2 * sp) - 1;
              const size_t k0 = (size_t)(so + 2 * u0);
              v -= ld_coh(sb + k0 * n + tc);
              if (j + sp < N) v -= ld_coh(sb + (k0 + 1) * n + tc);
              for (int k = lp; k < lp + 4 && k < l; ++k) so += 2 * (N >> (k + 1));
            }
          }
          if (g < 4) part[g * NP + tc] = v;
        }
        __syncthreads();
        if (t < NP) Y[t] = t < n ? xj[t] + ((part[t] + part[NP + t]) + (part[2 * NP + t] + part[3 * NP + t])) : T(0);
        __syncthreads();
        // y_j <- L_j^{-1} y_j:  out[c] = y[c] / L[c][c] + sum_{r<c} U[r][c] y[r]
        prod_upper_t<T, NB>(Us, Y, part, t);
        __syncthreads();
        if (t < NP) {
          T v = T(0);
          if (t < n) {
            v = Y[t] / D[t];
#pragma unroll
            for (int h = 0; h < S::H; ++h) v += part[h * NP + t];
            if (!ok) v = __builtin_nan("");
            st_coh(xj + t, v);
          }
          Y2[t] = v;
        }
        __syncthreads();
        T* sb = slots + (size_t)b * nslot * n;
        if (hasR) {
          T* dst = sb + (size_t)(sofl + 2 * tt) * n;
          prod_rl<T, NB>(r, Y2, t, [&](int row, T v) {
            if (row < n) st_coh(dst + row, v);
          });
        }
        asm volatile("" ::: "memory");  // one product's LDS operands in registers at a time
        if (hasC) {
          prod_cl<T, NB>(c, Y2, part2, t);
          __syncthreads();
          T* dst = sb + (size_t)(sofl + 2 * (tt - 1) + 1) * n;
          if (t < n) {
            T v = T(0);
#pragma unroll
            for (int h = 0; h < S::H; ++h) v += part2[h * NP + t];
            st_coh(dst + t, v);
          }
        }
        // publish: every thread's stores complete, then the counters
        stores_done();
        __syncthreads();
        if (t == 0) {
          if (hasR) __hip_atomic_fetch_add(fcount + (size_t)b * N + j + s, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (hasC) __hip_atomic_fetch_add(fcount + (size_t)b * N + j - s, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(fdone + (size_t)b * N + j, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      } else {
        if (t < NP) {
          Y[t] = t < n ? ld_coh(xj + t) : T(0);
          XA[t] = (hasR && t < n) ? ld_coh(x + ((size_t)b * N + j + s) * n + t) : T(0);
          XC[t] = (hasC && t < n) ? ld_coh(x + ((size_t)b * N + j - s) * n + t) : T(0);
        }
        __syncthreads();
        // z = y - R^T x_{j+s} - Lc x_{j-s}
        if (hasR) prod_cl<T, NB>(r, XA, part, t);
        asm volatile("" ::: "memory");
        if (hasC)
          prod_rl<T, NB>(c, XC, t, [&](int row, T v) { RW[row] = v; });
        __syncthreads();
        if (t < NP) {
          T v = T(0);
          if (t < n) {
            v = Y[t];
            if (hasR) {
#pragma unroll
              for (int h = 0; h < S::H; ++h) v -= part[h * NP + t];
            }
            if (hasC) v -= RW[t];
          }
          Y2[t] = v;
        }
        __syncthreads();
        // x_j <- L_j^{-T} z:  out[r] = z[r] / L[r][r] + sum_{k>r} U[r][k] z[k]
        prod_upper<T, NB>(Us, Y2, part2, t);
        __syncthreads();
        if (t < n) {
          T v = Y2[t] / D[t];
#pragma unroll
          for (int h = 0; h < S::H; ++h) v += part2[h * NP + t];
          if (!ok) v = __builtin_nan("");
          st_coh(xj + t, v);
        }
        stores_done();
        __syncthreads();
        if (t == 0) __hip_atomic_store(bdone + (size_t)b * N + j, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    __syncthreads();
  }
}

// Zeroes the ticket, status and per-node counters of the workspace.
__global__ __launch_bounds__(kThreads) void k_sweep_clear(int* __restrict__ ws, int count) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < count; i += gridDim.x * kThreads) ws[i] = 0;
}

}  // namespace

// Polls per dependency wait of the kernels enqueued from now on (a graph
// keeps the value it was captured with).
static std::atomic<int> g_spin_limit{kSweepSpinLimit};
void set_sweep_spin_limit(int polls) { g_spin_limit.store(polls == 0 ? kSweepSpinLimit : polls); }
int sweep_spin_limit() { return g_spin_limit.load(); }

namespace {

// Resident workgroups per CU of each instantiation, and the CU count.
int g_sweep_occ[2][4];
int g_cus = 0;

template <typename T, int NB>
hipError_t sweep_occupancy(int& out) {
  int nb = 0;
  const size_t lds = sweep_lds_bytes<T, NB>();
  hipError_t e = hipFuncSetAttribute((const void*)k_sweep<T, NB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep<T, NB>, SweepShape<T, NB>::THREADS, lds);
  out = nb > 0 ? nb : 1;
  return e;
}

template <typename T, int NB>
hipError_t launch_sweep(const T* L, const T* E, T* x, const uint8_t* mask, int* ws, T* slots, int* hstatus, int B,
                        int N, int n, int dirs, int lmin, int grid, hipStream_t st) {
  const size_t lds = sweep_lds_bytes<T, NB>();
  constexpr int threads = SweepShape<T, NB>::THREADS;
  hipLaunchKernelGGL((k_sweep<T, NB>), dim3(grid), dim3(threads), lds, st, L, E, x, mask, ws, slots, hstatus,
                     sweep_spin_limit(), B, N, n, dirs, lmin);
  return hipGetLastError();
}

}  // namespace

hipError_t init_sweep_kernels() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  if ((e = sweep_occupancy<float, 1>(g_sweep_occ[0][0])) != hipSuccess) return e;
  if ((e = sweep_occupancy<float, 2>(g_sweep_occ[0][1])) != hipSuccess) return e;
  if ((e = sweep_occupancy<float, 3>(g_sweep_occ[0][2])) != hipSuccess) return e;
  if ((e = sweep_occupancy<float, 4>(g_sweep_occ[0][3])) != hipSuccess) return e;
  if ((e = sweep_occupancy<double, 1>(g_sweep_occ[1][0])) != hipSuccess) return e;
  if ((e = sweep_occupancy<double, 2>(g_sweep_occ[1][1])) != hipSuccess) return e;
  if ((e = sweep_occupancy<double, 3>(g_sweep_occ[1][2])) != hipSuccess) return e;
  return sweep_occupancy<double, 4>(g_sweep_occ[1][3]);
}

// Measured on MI355X (solve of one right-hand side after the factor, N=512):
// the persistent upper-level launch beats the level kernels for FP32 blocks
// with n > 96 (205 vs 208 us per solve, and a converged refinement iteration
// costs two launches instead of ~40), but loses for smaller blocks and for
// FP64 (n=64 FP32: 275 vs 133 us; n=128 FP64: 379 vs 339 us): a dependency
// hop between workgroups costs about as much as a dependent kernel launch.
bool sweep_persistent_enabled(int n, int dsize) {
  if (n > kMaxN) return false;  // its node tasks hold whole blocks in LDS
  const char* v = getenv("SOCU_SWEEP");
  if (v != nullptr && strcmp(v, "level") == 0) return false;
  if (v != nullptr && strcmp(v, "persistent") == 0) return true;
  (void)dsize;
  return true;  // faster than the level kernels for every block size measured (see sweep_lmin)
}

size_t sweep_ws_ints(int B, int N) { return (size_t)kSweepHeader + 3 * (size_t)B * N; }

size_t sweep_slot_elems(int B, int N, int n) {
  const int k = sweep_slots_per_system(N);
  return (size_t)B * (size_t)(k > 0 ? k : 1) * n;
}

// Lowest level swept by the persistent launch: levels pay there while the
// widest of them has at most W nodes over the batch, with W = 64 for n <= 64,
// 256 for FP32 and 1024 for FP64 blocks above that.  Measured solve times
// (N=512, one right-hand side, lmin 0..5 against the level kernels) for
// B = 1, 4 and n = 32, 64, 96, 128 in both precisions: this choice is the
// fastest or within 4% of it in each case (FP64 n=128 B=1: 226 us against
// 341 us with level kernels only, FP32: 127 us against 220 us).  Below lmin
// the level kernels stream the wide levels.
int sweep_lmin(int B, int N, int n, int dsize) {
  const int Lv = recursive_iterations(N);
  const long long W = n <= 64 ? 64 : (dsize == 8 ? 1024 : 256);
  int lmin = 0;
  while (lmin < Lv - 1 && (long long)B * ((N + (1 << lmin)) >> (lmin + 1)) > W) ++lmin;
  return lmin;
}

template <typename T>
hipError_t sweep_persistent(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, bool fwd, bool bwd,
                            int* ws, T* slots, int* hstatus, hipStream_t st, bool clear) {
  if (B <= 0 || (!fwd && !bwd)) return hipSuccess;
  const int Lv = recursive_iterations(N);
  const int lmin = sweep_lmin(B, N, n, (int)sizeof(T));
  hipError_t e = hipSuccess;
  // forward below lmin: applies of levels 0..lmin-1 and their updates into
  // every receiving block (the last one without the next level's apply)
  if (fwd)
    for (int step = -1; step < lmin - 1 && e == hipSuccess; ++step)
      e = forward_step_blocked<T>(L, E, x, mask, B, N, n, 1, step, st);
  if (e == hipSuccess && fwd && lmin >= 1) e = forward_step_blocked<T>(L, E, x, mask, B, N, n, 1, lmin - 1, st, false);
  if (e != hipSuccess) return e;

  int total = 0;
  for (int l = lmin; l < Lv; ++l) total += ((N + (1 << l)) >> (l + 1)) * B;
  if (fwd && bwd) total *= 2;
  const int nb = (n + 31) / 32;
  const int occ = g_sweep_occ[sizeof(T) == 8 ? 1 : 0][nb - 1];
  int grid = (g_cus > 0 ? g_cus : 1) * occ;
  if (grid > total) grid = total;
  const int dirs = (fwd ? 1 : 0) | (bwd ? 2 : 0);
  const int nclear = (int)sweep_ws_ints(B, N);
  const int cblocks = (nclear + kThreads - 1) / kThreads;
  kt_begin(fwd ? kKcForward : kKcBackward, st);
  if (clear)
    hipLaunchKernelGGL(k_sweep_clear, dim3(cblocks < 1024 ? cblocks : 1024), dim3(kThreads), 0, st, ws, nclear);
  switch (nb) {
    case 1: e = launch_sweep<T, 1>(L, E, x, mask, ws, slots, hstatus, B, N, n, dirs, lmin, grid, st); break;
    case 2: e = launch_sweep<T, 2>(L, E, x, mask, ws, slots, hstatus, B, N, n, dirs, lmin, grid, st); break;
    case 3: e = launch_sweep<T, 3>(L, E, x, mask, ws, slots, hstatus, B, N, n, dirs, lmin, grid, st); break;
    case 4: e = launch_sweep<T, 4>(L, E, x, mask, ws, slots, hstatus, B, N, n, dirs, lmin, grid, st); break;
    default: e = hipErrorInvalidValue; break;
  }
  kt_end(fwd ? kKcForward : kKcBackward, st);
  if (e != hipSuccess) return e;
  // backward below lmin
  if (bwd)
    for (int l = lmin - 1; l >= 0 && e == hipSuccess; --l) e = backward_level_blocked<T>(L, E, x, mask, B, N, n, 1, l, st);
  return e;
}

template hipError_t sweep_persistent<float>(const float*, const float*, float*, const uint8_t*, int, int, int, bool,
                                            bool, int*, float*, int*, hipStream_t, bool);
template hipError_t sweep_persistent<double>(const double*, const double*, double*, const uint8_t*, int, int, int,
                                             bool, bool, int*, double*, int*, hipStream_t, bool);

}  // namespace socu
