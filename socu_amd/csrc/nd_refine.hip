// Mixed-precision factor_and_solve with FP64 iterative refinement.
//
// The reference has no mixed-precision path; this is an extension with the
// same data layout.  Psi (L, E in FP64) is left untouched; the plan owns a
// reduced-precision working copy (L32, E32) that is factored with the FP32
// kernels (Schur updates on FP32 MFMA).  Then, with x0 = 0 and r0 = b:
//   repeat:  d = Lp32^{-T} Lp32^{-1} r    (FP32 sweeps on the FP32 factor)
//            x += d                        (FP64)
//            r  = b - Psi x                (FP64, block-tridiagonal matvec)
//            stop when ||r||_F <= tol ||b||_F  (per batch entry)
// The stopping test runs on the device: every kernel of an iteration reads
// the per-system "done" flag and returns at once for converged systems, so
// the whole loop (fixed maximum iteration count) can be captured in one
// hipGraph and needs no host synchronisation.
#include "nd_common.h"
#include "nd_internal.h"

namespace socu {

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------

// FP32 working copy of Psi in one launch: per batch entry
//   tril(L32) <- tril(L) (lc elements), E32[0, ec) <- E (the tridiagonal
//   blocks); es is the per-entry stride of E and E32.  The fill-in region
//   E32[ec, es) is not touched: the factor writes each of its blocks (once,
//   as F = -R Lc) before anything reads it.
// The strict upper triangle of L32 is left as it is: the factor kernels read
// only tril of a diagonal block and then store its inverse there (the
// reference also reads only tril(L), potrf.py:101-103), so half of L is
// neither read nor written.
// VEC: four elements per step (two 16-byte loads, one 16-byte store), for
// lc, ec, es multiples of 4 and 16-byte aligned arrays; a step is skipped
// when its four elements lie in one row, strictly above the diagonal.
template <bool VEC>
__global__ __launch_bounds__(kThreads) void k_psi_to_f32(float* __restrict__ L32, float* __restrict__ E32,
                                                        const double* __restrict__ L, const double* __restrict__ E,
                                                        size_t lc, size_t ec, size_t es, int n,
                                                        const uint8_t* __restrict__ mask) {
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  constexpr int W = VEC ? 4 : 1;
  const size_t nl = lc / W, total = (lc + ec) / W;
  const double* Lb = L + (size_t)b * lc;
  const double* Eb = E + (size_t)b * es;
  float* L32b = L32 + (size_t)b * lc;
  float* E32b = E32 + (size_t)b * es;
  for (size_t u = (size_t)blockIdx.x * kThreads + threadIdx.x; u < total; u += (size_t)gridDim.x * kThreads) {
    if (VEC) {
      if (u < nl) {
        const int e = (int)((unsigned)(4 * u) % (unsigned)(n * n)), r = e / n, k = e - r * n;
        if (k > r && k + 3 < n) continue;
        const double2 a = reinterpret_cast<const double2*>(Lb)[2 * u];
        const double2 c = reinterpret_cast<const double2*>(Lb)[2 * u + 1];
        reinterpret_cast<float4*>(L32b)[u] = make_float4((float)a.x, (float)a.y, (float)c.x, (float)c.y);
      } else {
        const size_t v = u - nl;
        const double2 a = reinterpret_cast<const double2*>(Eb)[2 * v];
        const double2 c = reinterpret_cast<const double2*>(Eb)[2 * v + 1];
        reinterpret_cast<float4*>(E32b)[v] = make_float4((float)a.x, (float)a.y, (float)c.x, (float)c.y);
      }
    } else {
      if (u < nl) {
        const int e = (int)((unsigned)u % (unsigned)(n * n)), r = e / n;
        if (e - r * n <= r) L32b[u] = (float)Lb[u];
      }
      else E32b[u - nl] = (float)Eb[u - nl];
    }
  }
}

// Refinement state per batch entry (4 doubles): [0] = ||b||^2, [1] = ||r||^2
// of the last residual, [2] = done flag (1 when converged or masked out),
// [3] = correction solves applied.  active[b] = !done, as a byte mask for the
// FP32 sweep kernels.
__device__ __forceinline__ bool refine_done(const double* state, int b) { return state[4 * b + 2] != 0.0; }

// x (FP64) <- 0 and r32 <- b (FP32); records ||b||^2 partials.
__global__ __launch_bounds__(kThreads) void k_refine_init(double* __restrict__ x, const double* __restrict__ b_in,
                                                         float* __restrict__ r32, double* __restrict__ partial,
                                                         size_t count, int nblk, const uint8_t* __restrict__ mask) {
  const int b = blockIdx.y;
  __shared__ double red[kThreads];
  double acc = 0.0;
  if (!masked_out(mask, b)) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < count; i += (size_t)gridDim.x * kThreads) {
      const double v = b_in[(size_t)b * count + i];
      r32[(size_t)b * count + i] = (float)v;
      x[(size_t)b * count + i] = 0.0;
      acc += v * v;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)b * nblk + blockIdx.x] = red[0];
}

// Sum of partial[b][0..nblk) by one workgroup (fixed order: deterministic).
__device__ __forceinline__ double block_sum(const double* __restrict__ part, int nblk, double* red) {
  double s = 0.0;
  for (int q = threadIdx.x; q < nblk; q += kThreads) s += part[q];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  return red[0];
}

// Re-arms system b's part of the persistent sweep workspace for the next
// sweep (the solve's sweeps then need no clearing launch): its per-node
// counters, and from workgroup 0 the ticket and, with status, the status word
// (zeroed once per solve, so that an expired wait stays visible to the host).
__device__ __forceinline__ void sweep_ws_zero(int* __restrict__ ws, int B, int N, int b, bool status) {
  if (ws == nullptr) return;
  int* cnt = ws + kSweepHeader;
  for (int k = threadIdx.x; k < 3 * N; k += kThreads) {
    const int a = k / N;
    cnt[(size_t)a * B * N + (size_t)b * N + (k - a * N)] = 0;
  }
  if (b == 0 && threadIdx.x == 0) {
    ws[0] = 0;
    if (status) ws[1] = 0;
  }
}

// state[b] <- (||b||^2, ||b||^2, done = masked out, 0); one workgroup per b.
// Also zeroes ticket[b] (the one-RHS residual's count of finished workgroups).
__global__ __launch_bounds__(kThreads) void k_refine_start(double* __restrict__ state, uint8_t* __restrict__ active,
                                                          const double* __restrict__ partial, int nblk,
                                                          const uint8_t* __restrict__ mask, int* __restrict__ sweep_ws,
                                                          int* __restrict__ ticket, int N) {
  const int b = blockIdx.x;
  sweep_ws_zero(sweep_ws, gridDim.x, N, b, true);
  if (ticket != nullptr && threadIdx.x == 0) ticket[b] = 0;
  __shared__ double red[kThreads];
  const double s = block_sum(partial + (size_t)b * nblk, nblk, red);
  if (threadIdx.x == 0) {
    const bool off = masked_out(mask, b);
    state[4 * b + 0] = s;
    state[4 * b + 1] = s;
    state[4 * b + 2] = off ? 1.0 : 0.0;
    state[4 * b + 3] = 0.0;
    active[b] = off ? 0 : 1;
  }
}

// x += d (FP32 correction, converted), for systems not yet converged.
__global__ __launch_bounds__(kThreads) void k_refine_update(double* __restrict__ x, const float* __restrict__ d,
                                                           size_t count, const double* __restrict__ state) {
  const int b = blockIdx.y;
  if (refine_done(state, b)) return;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < count; i += (size_t)gridDim.x * kThreads)
    x[(size_t)b * count + i] += (double)d[(size_t)b * count + i];
}

// r = b - Psi x (FP64) for a slice of kResRows rows of block row i, one
// workgroup per (block row, slice, batch):
//   r_i = b_i - sym(L_i) x_i - E_{i-1} x_{i-1} - E_i^T x_{i+1}
// Only tril(L_i) is read (the matrix the reference's solver sees).
// Row-access products (tril(L_i) x_i, E_{i-1} x_{i-1}) use 8 lanes per row
// and a shuffle reduction; column-access products (strict upper of sym(L_i)
// from the lower triangle, E_i^T x_{i+1}) use 8 groups of 32 lanes, lane r of
// a group owning output row r, so that each load instruction reads 32
// consecutive entries of one matrix row, and the 8 partial sums meet in LDS.
// The FP32 copy of r feeds the next correction solve; partial[b][blockIdx.x]
// gets the slice's ||r||^2.
constexpr int kResM = 8;      // right-hand sides per pass
constexpr int kResRows = 32;  // output rows per workgroup
__global__ __launch_bounds__(kThreads) void k_residual(const double* __restrict__ L, const double* __restrict__ E,
                                                      const double* __restrict__ x, const double* __restrict__ b_in,
                                                      float* __restrict__ r32, double* __restrict__ partial,
                                                      const double* __restrict__ state, int N, int n, int m, int nE,
                                                      int nsl) {
  const int b = blockIdx.y;
  const int i = blockIdx.x / nsl, r0 = (blockIdx.x - i * nsl) * kResRows;
  if (refine_done(state, b)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  double* xs = reinterpret_cast<double*>(smem_raw);  // 3 x n x m: x_{i-1}, x_i, x_{i+1}
  __shared__ double part[8][kResRows][kResM];
  __shared__ double red[kThreads];
  const double* xb = x + (size_t)b * N * nm;
  for (int q = threadIdx.x; q < 3 * (int)nm; q += kThreads) {
    const int which = q / (int)nm, e = q - which * (int)nm;
    const int row = i - 1 + which;
    xs[q] = (row >= 0 && row < N) ? xb[(size_t)row * nm + e] : 0.0;
  }
  __syncthreads();
  const double* Li = L + ((size_t)b * N + i) * nn;
  const double* Eb = E + (size_t)b * nE * nn;
  const double* Em = (i > 0) ? Eb + (size_t)(i - 1) * nn : nullptr;  // block (i, i-1)
  const double* Ep = (i + 1 < N) ? Eb + (size_t)i * nn : nullptr;    // block (i+1, i)
  const double* x_prev = xs;
  const double* x_cur = xs + nm;
  const double* x_next = xs + 2 * nm;
  const double* bi = b_in + (size_t)b * N * nm + (size_t)i * nm;
  float* ri = r32 + (size_t)b * N * nm + (size_t)i * nm;
  const int tid = threadIdx.x;
  const int cr = tid & 31, cg = tid >> 5;  // column products: row r0 + cr, k = cg (mod 8)
  const int rr = tid >> 3, rs = tid & 7;   // row products: row r0 + rr, k = rs (mod 8)
  double sq = 0.0;
  for (int c0 = 0; c0 < m; c0 += kResM) {
    const int w = min(kResM, m - c0);
    {
      const int r = r0 + cr;
      double acc[kResM];
#pragma unroll
      for (int c = 0; c < kResM; ++c) acc[c] = 0.0;
      if (r < n) {
        // chunks of 4 inner indices: unconditional loads (clamped), pinned
        // after the chunk is issued, masked afterwards
        const double* Epr = Ep ? Ep : Li;
        for (int kb = cg; kb < n; kb += 32) {
          double lv[4], ev[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int k = min(kb + 8 * u, n - 1);
            lv[u] = Li[(size_t)k * n + r];
            ev[u] = Epr[(size_t)k * n + r];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            pin_reg(lv[u]);
            pin_reg(ev[u]);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int k = kb + 8 * u;
            if (k < n) {
              const double l1 = (k > r) ? lv[u] : 0.0, e1 = Ep ? ev[u] : 0.0;
#pragma unroll
              for (int c = 0; c < kResM; ++c)
                if (c < w) acc[c] += l1 * x_cur[(size_t)k * m + c0 + c] + e1 * x_next[(size_t)k * m + c0 + c];
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < kResM; ++c) part[cg][cr][c] = acc[c];
    }
    double racc[kResM];
#pragma unroll
    for (int c = 0; c < kResM; ++c) racc[c] = 0.0;
    {
      const int r = r0 + rr;
      if (r < n) {
        const double* Emr = Em ? Em : Li;
        for (int kb = rs; kb < n; kb += 32) {
          double lv[4], ev[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int k = min(kb + 8 * u, n - 1);
            lv[u] = Li[(size_t)r * n + k];
            ev[u] = Emr[(size_t)r * n + k];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            pin_reg(lv[u]);
            pin_reg(ev[u]);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int k = kb + 8 * u;
            if (k < n) {
              const double l1 = (k <= r) ? lv[u] : 0.0, e1 = Em ? ev[u] : 0.0;
#pragma unroll
              for (int c = 0; c < kResM; ++c)
                if (c < w) racc[c] += l1 * x_cur[(size_t)k * m + c0 + c] + e1 * x_prev[(size_t)k * m + c0 + c];
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < kResM; ++c)
        for (int o = 4; o >= 1; o >>= 1) racc[c] += __shfl_xor(racc[c], o, 8);
    }
    __syncthreads();
    if (rs == 0 && r0 + rr < n) {
      const int r = r0 + rr;
      for (int c = 0; c < w; ++c) {
        double v = racc[c];
#pragma unroll
        for (int g = 0; g < 8; ++g) v += part[g][rr][c];
        const double res = bi[(size_t)r * m + c0 + c] - v;
        ri[(size_t)r * m + c0 + c] = (float)res;
        sq += res * res;
      }
    }
    __syncthreads();
  }
  red[tid] = sq;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  if (tid == 0) partial[(size_t)b * gridDim.x + blockIdx.x] = red[0];
}

// ||r||^2 from the partials; sets the done flag when ||r|| <= tol ||b||.
// One workgroup per system.
__global__ __launch_bounds__(kThreads) void k_refine_check(double* __restrict__ state, uint8_t* __restrict__ active,
                                                          const double* __restrict__ partial, int nblk, double tol2,
                                                          int* __restrict__ sweep_ws, int N) {
  const int b = blockIdx.x;
  sweep_ws_zero(sweep_ws, gridDim.x, N, b, false);
  if (state[4 * b + 2] != 0.0) return;
  __shared__ double red[kThreads];
  const double s = block_sum(partial + (size_t)b * nblk, nblk, red);
  if (threadIdx.x == 0) {
    state[4 * b + 1] = s;
    state[4 * b + 3] += 1.0;
    if (s <= tol2 * state[4 * b + 0] || !(s == s)) {  // converged, or NaN (non-SPD input)
      state[4 * b + 2] = 1.0;
      active[b] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// One right-hand side (m == 1): update, residual and convergence test in two
// launches that read tril(L_i) and E_i once each (the general kernel above
// reads both triangles of L_i and every E block twice).  With y = x + d:
//   k_resid1_blocks, one workgroup per (block row i, slice s of 32 rows):
//     rows r of the slice:  W_i[r]     = b_i[r] - sum_{k<=r} L_i[r][k] y_i[k]
//                           V_{i+1}[r] = sum_k E_i[r][k] y_i[k]   (E_i = block (i+1, i))
//     every column k, summed over the slice's rows:
//                           C_{i,s}[k] = sum_{r>k} L_i[r][k] y_i[r] + sum_r E_i[r][k] y_{i+1}[r]
//   k_resid1_finish: x += d, r_i = W_i - sum_s C_{i,s} - V_i, r32 <- r and the
//     ||r||^2 partials; the system's last workgroup to finish (a relaxed
//     counter; partials pass through agent-scope atomic stores and loads, so
//     no L2 write-back fence is needed) sums them in a fixed order and
//     applies the convergence test.
// Lanes: 16 per row (each load instruction reads 128 contiguous bytes of four
// rows), row r = 32 s + (tid >> 4) + 16 g, column k = (tid & 15) + 16 j.
__global__ __launch_bounds__(kThreads) void k_resid1_blocks(const double* __restrict__ L, const double* __restrict__ E,
                                                           const double* __restrict__ x, const float* __restrict__ d,
                                                           const double* __restrict__ b_in, double* __restrict__ W,
                                                           double* __restrict__ V, double* __restrict__ C,
                                                           const double* __restrict__ state, int N, int n, int nE,
                                                           int nsl) {
  const int b = blockIdx.y;
  if (refine_done(state, b)) return;
  const int i = blockIdx.x / nsl, s = blockIdx.x - i * nsl;
  const size_t nn = (size_t)n * n;
  __shared__ double yc[kMaxN], yn[kMaxN];
  __shared__ double redc[16][kMaxN];
  __shared__ double redr[2][32][17];
  const double* xb = x + (size_t)b * N * n;
  const float* db = d + (size_t)b * N * n;
  const bool hasE = i + 1 < N;
  for (int q = threadIdx.x; q < 2 * n; q += kThreads) {
    const int w = q >= n ? 1 : 0, e = q - w * n, row = i + w;
    const double y = (row < N) ? xb[(size_t)row * n + e] + (double)db[(size_t)row * n + e] : 0.0;
    (w ? yn : yc)[e] = y;
  }
  const int t = threadIdx.x, rr = t >> 4, rs = t & 15;
  const double* Li = L + ((size_t)b * N + i) * nn;
  const double* Ei = E + (size_t)b * nE * nn + (size_t)i * nn;  // read only when hasE
  double ev[2][8], lv[2][8];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const int r = 32 * s + rr + 16 * g;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = rs + 16 * j;
      const bool in = r < n && k < n;
      ev[g][j] = (hasE && in) ? Ei[(size_t)r * n + k] : 0.0;
      lv[g][j] = (in && 16 * j <= r) ? Li[(size_t)r * n + k] : 0.0;  // skips segments above the diagonal
    }
  }
  __syncthreads();  // yc, yn
  double ycol[8], cacc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = rs + 16 * j;
    ycol[j] = (k < n) ? yc[k] : 0.0;
    cacc[j] = 0.0;
  }
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const int q = rr + 16 * g, r = 32 * s + q;
    double ra = 0.0, va = 0.0;
    if (r < n) {
      const double yr = yc[r], ynr = yn[r];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = rs + 16 * j;
        va = fma(ev[g][j], ycol[j], va);
        cacc[j] = fma(ev[g][j], ynr, cacc[j]);
        ra = fma((k <= r) ? lv[g][j] : 0.0, ycol[j], ra);
        cacc[j] = fma((k < r) ? lv[g][j] : 0.0, yr, cacc[j]);
      }
    }
    redr[0][q][rs] = ra;
    redr[1][q][rs] = va;
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) redc[rr][rs + 16 * j] = cacc[j];
  __syncthreads();
  const size_t row0 = ((size_t)b * N + i) * n;
  if (t < 64) {
    const int which = t >> 5, q = t & 31, r = 32 * s + q;
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) sum += redr[which][q][u];
    if (r < n) {
      if (which == 0) W[row0 + r] = b_in[row0 + r] - sum;
      else if (hasE) V[row0 + n + r] = sum;
    }
  }
  for (int k = t; k < n; k += kThreads) {
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) sum += redc[u][k];
    C[(((size_t)b * N + i) * nsl + s) * n + k] = sum;
  }
}

// dr holds the correction d on entry and the FP32 residual on exit.  The
// workgroups of all systems also re-arm the persistent sweep workspace, as
// k_refine_check does.
__global__ __launch_bounds__(kThreads) void k_resid1_finish(double* __restrict__ x, float* __restrict__ dr,
                                                           const double* __restrict__ W, const double* __restrict__ V,
                                                           const double* __restrict__ C, double* __restrict__ partial,
                                                           double* __restrict__ state, uint8_t* __restrict__ active,
                                                           int* __restrict__ ticket, double tol2,
                                                           int* __restrict__ sweep_ws, int N, int n, int nsl) {
  const int b = blockIdx.y;
  if (sweep_ws != nullptr) {
    int* cnt = sweep_ws + kSweepHeader;
    const int B = gridDim.y;
    for (int k = blockIdx.x * kThreads + threadIdx.x; k < 3 * N; k += gridDim.x * kThreads) {
      const int a = k / N;
      cnt[(size_t)a * B * N + (size_t)b * N + (k - a * N)] = 0;
    }
    if (b == 0 && blockIdx.x == 0 && threadIdx.x == 0) sweep_ws[0] = 0;
  }
  if (refine_done(state, b)) return;
  const int count = N * n;  // < 2^31 (checked by the host)
  __shared__ double red[kThreads / 64];
  const size_t base = (size_t)b * count;
  double sq = 0.0;
  for (int e = blockIdx.x * kThreads + threadIdx.x; e < count; e += gridDim.x * kThreads) {
    const size_t gi = base + e;
    const int i = e / n, a = e - i * n;
    double r = W[gi];
    const double* Ci = C + (((size_t)b * N + i) * nsl) * n + a;
    for (int q = 0; q < nsl; ++q) r -= Ci[q * n];
    if (i > 0) r -= V[gi];
    x[gi] += (double)dr[gi];
    dr[gi] = (float)r;
    sq += r * r;
  }
  // wavefront sums, then the four of the workgroup in a fixed order
  for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  double* part = partial + (size_t)b * gridDim.x;
  __shared__ int last;
  if (threadIdx.x == 0) {
    st_coh(part + blockIdx.x, (red[0] + red[1]) + (red[2] + red[3]));
    stores_done();
    const int done = __hip_atomic_fetch_add(ticket + b, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = done == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  // the system's last workgroup: sum the partials as k_refine_check does
  __shared__ double tot[kThreads];
  double s = 0.0;
  for (int q = threadIdx.x; q < (int)gridDim.x; q += kThreads) s += ld_coh(part + q);
  tot[threadIdx.x] = s;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) tot[threadIdx.x] += tot[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double s2 = tot[0];
    ticket[b] = 0;
    state[4 * b + 1] = s2;
    state[4 * b + 3] += 1.0;
    if (s2 <= tol2 * state[4 * b + 0] || !(s2 == s2)) {  // converged, or NaN (non-SPD input)
      state[4 * b + 2] = 1.0;
      active[b] = 0;
    }
  }
}

}  // namespace socu

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
namespace socu {

static int refine_blocks(size_t count) {
  const size_t blk = (count + kThreads - 1) / kThreads;
  return (int)(blk < 256 ? (blk > 0 ? blk : 1) : 256);
}

size_t mixed_residual_lds(int n, int m) { return sizeof(double) * 3 * (size_t)n * m; }

// Dynamic LDS cap of the residual kernel (its static LDS is 18 KiB of the
// 160 KiB per workgroup).
constexpr size_t kResidualLdsMax = 140 * 1024;

static int residual_slices(int n) { return (n + kResRows - 1) / kResRows; }


hipError_t init_refine_kernels() {
  return hipFuncSetAttribute((const void*)k_residual, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kResidualLdsMax);
}

bool mixed_fits(int n, int m) { return mixed_residual_lds(n, m) <= kResidualLdsMax; }

hipError_t mixed_alloc(MixedPlan& p) {
  const size_t nn = (size_t)p.n * p.n, nm = (size_t)p.n * p.m;
  p.nE = (int)off_diag_storage_len(p.N);
  p.nblk_init = refine_blocks((size_t)p.N * nm);
  hipError_t e;
  if ((e = hipMalloc(&p.L32, sizeof(float) * p.B * (size_t)p.N * nn)) != hipSuccess) return e;
  if ((e = hipMalloc(&p.E32, sizeof(float) * p.B * (size_t)(p.nE > 0 ? p.nE : 1) * nn)) != hipSuccess) return e;
  if ((e = hipMalloc(&p.d32, sizeof(float) * p.B * (size_t)p.N * nm)) != hipSuccess) return e;
  if ((e = hipMalloc(&p.b64, sizeof(double) * p.B * (size_t)p.N * nm)) != hipSuccess) return e;
  if ((e = hipMalloc(&p.state, sizeof(double) * 4 * p.B)) != hipSuccess) return e;
  const int nres = p.N * residual_slices(p.n);
  const int np = nres > p.nblk_init ? nres : p.nblk_init;
  if ((e = hipMalloc(&p.partial, sizeof(double) * p.B * (size_t)np)) != hipSuccess) return e;
  if ((e = hipMalloc(&p.active, (size_t)p.B)) != hipSuccess) return e;
  if (p.m == 1) {
    const size_t rows = (size_t)p.B * p.N * p.n;
    if ((e = hipMalloc(&p.res_w, sizeof(double) * rows)) != hipSuccess) return e;
    if ((e = hipMalloc(&p.res_v, sizeof(double) * rows)) != hipSuccess) return e;
    if ((e = hipMalloc(&p.res_c, sizeof(double) * rows * residual_slices(p.n))) != hipSuccess) return e;
    if ((e = hipMalloc(&p.res_ticket, sizeof(int) * p.B)) != hipSuccess) return e;
  }
  if (p.m == 1 && sweep_persistent_enabled(p.n, (int)sizeof(float))) {
    if ((e = hipMalloc(&p.sweep_ws, sizeof(int) * sweep_ws_ints(p.B, p.N))) != hipSuccess) return e;
    if ((e = hipMemset(p.sweep_ws, 0, sizeof(int) * sweep_ws_ints(p.B, p.N))) != hipSuccess) return e;
    if ((e = hipMalloc(&p.sweep_slots, sizeof(float) * sweep_slot_elems(p.B, p.N, p.n))) != hipSuccess) return e;
    if ((e = hipHostMalloc(&p.sweep_hstatus, sizeof(int), hipHostMallocMapped)) != hipSuccess) {
      p.sweep_hstatus = nullptr;
      return e;
    }
    *p.sweep_hstatus = 0;
    if ((e = hipHostGetDevicePointer((void**)&p.sweep_hstatus_dev, p.sweep_hstatus, 0)) != hipSuccess) return e;
  }
  return p.use_overlap ? overlap_create(p.overlap) : hipSuccess;
}

void mixed_free(MixedPlan& p) {
  void* ptrs[] = {p.L32,    p.E32,         p.d32,   p.b64,   p.state, p.partial,  p.active,
                  p.sweep_ws, p.sweep_slots, p.res_w, p.res_v, p.res_c, p.res_ticket};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (p.sweep_hstatus) (void)hipHostFree(p.sweep_hstatus);
  p.sweep_hstatus = p.sweep_hstatus_dev = nullptr;
  p.L32 = p.E32 = p.d32 = nullptr;
  p.b64 = p.state = p.partial = nullptr;
  p.active = nullptr;
  p.sweep_ws = nullptr;
  p.sweep_slots = nullptr;
  p.res_w = p.res_v = p.res_c = nullptr;
  p.res_ticket = nullptr;
  overlap_destroy(p.overlap);
}

hipError_t mixed_enqueue(const MixedPlan& p, hipStream_t st) {
  const int B = p.B, N = p.N, n = p.n, m = p.m, nE = p.nE;
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  const size_t xcount = (size_t)N * nm;
  hipError_t e = hipMemcpyAsync(p.b64, p.x, sizeof(double) * B * xcount, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return e;
  // FP32 working copy of Psi (tril(L) and the N-1 coupling blocks of E; the
  // factor writes the fill-in region), in one or two launches: lcn elements
  // of L, ecn of E, escn the per-entry stride of E's storage.
  const size_t lc = (size_t)N * nn;
  const size_t ec = N > 1 ? (size_t)(N - 1) * nn : 0, es = N > 1 ? (size_t)nE * nn : 0;
  auto cast = [&](size_t lcn, size_t ecn, size_t escn, hipStream_t s) {
    const bool vec = lcn % 4 == 0 && ecn % 4 == 0 && escn % 4 == 0 && ((uintptr_t)p.L % 16) == 0 &&
                     ((uintptr_t)p.E % 16) == 0 && ((uintptr_t)p.L32 % 16) == 0 && ((uintptr_t)p.E32 % 16) == 0;
    const size_t units = (lcn + ecn) / (vec ? 4 : 1);
    size_t blocks = (units + kThreads - 1) / kThreads;
    const size_t cap = (size_t)(2048 / B > 0 ? 2048 / B : 1);
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    const double* Ep = N > 1 ? p.E : p.L;  // unused when escn == 0
    float* E32p = N > 1 ? p.E32 : p.L32;
    if (vec)
      hipLaunchKernelGGL((k_psi_to_f32<true>), dim3((unsigned)blocks, B), dim3(kThreads), 0, s, p.L32, E32p, p.L, Ep,
                         lcn, ecn, escn, n, p.mask);
    else
      hipLaunchKernelGGL((k_psi_to_f32<false>), dim3((unsigned)blocks, B), dim3(kThreads), 0, s, p.L32, E32p, p.L,
                         Ep, lcn, ecn, escn, n, p.mask);
  };
  // x <- 0 and the first right-hand side d32 <- b: independent of the factor
  auto init = [&](hipStream_t s) {
    hipLaunchKernelGGL(k_refine_init, dim3(p.nblk_init, B), dim3(kThreads), 0, s, p.x, p.b64, p.d32, p.partial,
                       xcount, p.nblk_init, p.mask);
    hipLaunchKernelGGL(k_refine_start, dim3(B), dim3(kThreads), 0, s, p.state, p.active, p.partial, p.nblk_init,
                       p.mask, p.sweep_ws, p.res_ticket, N);
  };
  const int Lv = recursive_iterations(N);
  const bool overlap = p.overlap.side != nullptr;
  if (overlap) {
    // The factor's first kernel needs only tril(L); E, the refinement state
    // and the first correction solve's forward sweep go to the side stream,
    // alongside the factor (the first trsm waits for E).
    cast(lc, 0, 0, st);
    const std::function<hipError_t(hipStream_t)> first = [&](hipStream_t s) {
      cast(0, ec, es, s);
      init(s);
      return hipGetLastError();
    };
    if ((e = factor_with_forward<float>(p.L32, p.E32, p.d32, p.mask, p.active, B, N, n, m, p.prec, p.overlap,
                                        st, &first)) != hipSuccess)
      return e;
  } else {
    cast(lc, ec, es, st);
    init(st);
    for (int l = 0; l < Lv; ++l)
      if ((e = factor_level_blocked<float>(p.L32, p.E32, p.mask, B, N, n, l, p.prec, st)) != hipSuccess)
        return e;
  }
  const size_t rlds = mixed_residual_lds(n, m);
  for (int it = 0; it < p.max_iter; ++it) {
    const bool fwd = !(overlap && it == 0);  // done alongside the factor
    if (p.sweep_ws != nullptr) {
      // the upper levels of both sweeps in one persistent launch; converged
      // systems are skipped
      // (the workspace is re-armed by k_refine_start and by k_resid1_finish
      // or k_refine_check)
      if ((e = sweep_persistent<float>(p.L32, p.E32, p.d32, p.active, B, N, n, fwd, true, p.sweep_ws, p.sweep_slots,
                                       p.sweep_hstatus_dev, st, false)) != hipSuccess)
        return e;
    } else {
      if (fwd)
        for (int step = -1; step < Lv - 1; ++step)
          if ((e = forward_step_blocked<float>(p.L32, p.E32, p.d32, p.active, B, N, n, m, step, st)) != hipSuccess)
            return e;
      for (int l = Lv - 1; l >= 0; --l)
        if ((e = backward_level_blocked<float>(p.L32, p.E32, p.d32, p.active, B, N, n, m, l, st)) != hipSuccess)
          return e;
    }
    kt_begin(kKcRefine, st);
    if (p.res_w != nullptr) {
      const int nsl = residual_slices(n);
      hipLaunchKernelGGL(k_resid1_blocks, dim3(N * nsl, B), dim3(kThreads), 0, st, p.L, p.E, p.x, p.d32, p.b64,
                         p.res_w, p.res_v, p.res_c, p.state, N, n, nE, nsl);
      const int nb2 = refine_blocks(xcount);
      hipLaunchKernelGGL(k_resid1_finish, dim3(nb2, B), dim3(kThreads), 0, st, p.x, p.d32, p.res_w, p.res_v,
                         p.res_c, p.partial, p.state, p.active, p.res_ticket, p.tol * p.tol, p.sweep_ws, N, n, nsl);
      kt_end(kKcRefine, st);
      continue;
    }
    hipLaunchKernelGGL(k_refine_update, dim3(refine_blocks(xcount), B), dim3(kThreads), 0, st, p.x, p.d32, xcount,
                       p.state);
    const int nsl = residual_slices(n);
    hipLaunchKernelGGL(k_residual, dim3(N * nsl, B), dim3(kThreads), rlds, st, p.L, p.E, p.x, p.b64, p.d32,
                       p.partial, p.state, N, n, m, nE, nsl);
    hipLaunchKernelGGL(k_refine_check, dim3(B), dim3(kThreads), 0, st, p.state, p.active, p.partial, N * nsl,
                       p.tol * p.tol, p.sweep_ws, N);
    kt_end(kKcRefine, st);
  }
  return hipGetLastError();
}

}  // namespace socu
