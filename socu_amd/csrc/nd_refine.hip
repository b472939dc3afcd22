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

// dst (float) <- src (double) over count elements of each batch entry.
__global__ __launch_bounds__(kThreads) void k_cast_f64_f32(float* __restrict__ dst, const double* __restrict__ src,
                                                          size_t count, size_t dst_stride, size_t src_stride,
                                                          const uint8_t* __restrict__ mask) {
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  const double* s = src + (size_t)b * src_stride;
  float* d = dst + (size_t)b * dst_stride;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < count; i += (size_t)gridDim.x * kThreads)
    d[i] = (float)s[i];
}

// dst (float) <- 0 over count elements of each batch entry.
__global__ __launch_bounds__(kThreads) void k_zero_f32(float* __restrict__ dst, size_t count, size_t stride,
                                                      const uint8_t* __restrict__ mask) {
  const int b = blockIdx.y;
  if (masked_out(mask, b)) return;
  float* d = dst + (size_t)b * stride;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < count; i += (size_t)gridDim.x * kThreads)
    d[i] = 0.f;
}

// Refinement state per batch entry: [0] = ||b||^2, [1] = ||r||^2 of the last
// residual, [2] = done flag (1.0 when converged or masked out).
struct RefineState {
  double* state;     // B x 3
  double* partial;   // B x nblk partial sums of squares
  int nblk;          // residual workgroups per batch entry
};

__device__ __forceinline__ bool refine_done(const double* state, int b) { return state[3 * b + 2] != 0.0; }

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

// state[b] <- (sum partials, 0, done = masked_out); run with one thread per b.
__global__ void k_refine_start(double* __restrict__ state, const double* __restrict__ partial, int nblk, int B,
                               const uint8_t* __restrict__ mask) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double s = 0.0;
  for (int q = 0; q < nblk; ++q) s += partial[(size_t)b * nblk + q];
  state[3 * b + 0] = s;
  state[3 * b + 1] = s;
  state[3 * b + 2] = masked_out(mask, b) ? 1.0 : 0.0;
}

// x += d (FP32 correction, converted), for systems not yet converged.
__global__ __launch_bounds__(kThreads) void k_refine_update(double* __restrict__ x, const float* __restrict__ d,
                                                           size_t count, const double* __restrict__ state) {
  const int b = blockIdx.y;
  if (refine_done(state, b)) return;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < count; i += (size_t)gridDim.x * kThreads)
    x[(size_t)b * count + i] += (double)d[(size_t)b * count + i];
}

// r = b - Psi x for block row i (FP64), one workgroup per (row, batch):
//   r_i = b_i - sym(L_i) x_i - E_{i-1} x_{i-1} - E_i^T x_{i+1}
// Only tril(L_i) is read (the matrix the reference's solver sees).  The FP32
// copy r32 feeds the next correction solve; partial[b][i] gets ||r_i||^2.
template <int M_MAX>
__global__ __launch_bounds__(kThreads) void k_residual(const double* __restrict__ L, const double* __restrict__ E,
                                                      const double* __restrict__ x, const double* __restrict__ b_in,
                                                      float* __restrict__ r32, double* __restrict__ partial,
                                                      const double* __restrict__ state, int N, int n, int m, int nE) {
  const int b = blockIdx.y;
  const int i = blockIdx.x;
  if (refine_done(state, b)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* xs = reinterpret_cast<double*>(smem_raw);  // 3 x n x m: x_{i-1}, x_i, x_{i+1}
  __shared__ double red[kThreads];
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  const double* xb = x + (size_t)b * N * nm;
  for (int q = threadIdx.x; q < 3 * (int)nm; q += kThreads) {
    const int which = q / (int)nm, e = q - which * (int)nm;
    const int row = i - 1 + which;
    xs[q] = (row >= 0 && row < N) ? xb[(size_t)row * nm + e] : 0.0;
  }
  __syncthreads();
  const double* Li = L + ((size_t)b * N + i) * nn;
  const double* Eb = E + (size_t)b * nE * nn;
  double acc = 0.0;
  // thread -> (r, c) output; rows split across threads, each thread walks k.
  for (int q = threadIdx.x; q < (int)nm; q += kThreads) {
    const int r = q / m, c = q - r * m;
    double s = b_in[(size_t)b * N * nm + (size_t)i * nm + q];
    for (int k = 0; k < n; ++k) {
      const double lik = (k <= r) ? Li[(size_t)r * n + k] : Li[(size_t)k * n + r];
      s -= lik * xs[nm + (size_t)k * m + c];
    }
    if (i > 0) {
      const double* Em = Eb + (size_t)(i - 1) * nn;  // block (i, i-1)
      for (int k = 0; k < n; ++k) s -= Em[(size_t)r * n + k] * xs[(size_t)k * m + c];
    }
    if (i + 1 < N) {
      const double* Ep = Eb + (size_t)i * nn;  // block (i+1, i): contributes E_i^T x_{i+1}
      for (int k = 0; k < n; ++k) s -= Ep[(size_t)k * n + r] * xs[2 * nm + (size_t)k * m + c];
    }
    r32[(size_t)b * N * nm + (size_t)i * nm + q] = (float)s;
    acc += s * s;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)b * N + i] = red[0];
}

// ||r||^2 from the partials; sets the done flag when ||r|| <= tol ||b||.
__global__ void k_refine_check(double* __restrict__ state, const double* __restrict__ partial, int nblk, int B,
                               double tol2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  if (state[3 * b + 2] != 0.0) return;
  double s = 0.0;
  for (int q = 0; q < nblk; ++q) s += partial[(size_t)b * nblk + q];
  state[3 * b + 1] = s;
  if (s <= tol2 * state[3 * b + 0]) state[3 * b + 2] = 1.0;
}

}  // namespace socu
