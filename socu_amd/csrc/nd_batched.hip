// Batched ND factor / solve for many independent systems of uneven size
// (horizon N_p and block size n_p <= 32 per system), one persistent
// workgroup per system.  The reference can only batch systems that share N
// and n (block_tridiag_solver.py:749-768); this path has no counterpart there
// but computes the same factor, in the same storage, per system.
//
// A workgroup walks all ND levels of its system.  Within a level, the waves
// share the node tasks (potrf + inverse, both trsm, y_i <- Linv y_i), a
// barrier, then the receiving-block tasks (L_j -= R R^T + Lc^T Lc, fill-in
// F -= R Lc, y_j -= R y_{j-s} + Lc^T y_{j+s}), and a barrier.  The backward
// sweep walks the levels down with one task per node.  No two tasks of a
// phase write the same block, so there are no atomics and no inter-workgroup
// synchronisation.  Tiles are padded to NT in {8, 16, 32}; a wavefront runs
// 64/(2*NT) node factorisations at once (tile_chol_inv_grouped).  Systems are
// dispatched in order of decreasing cost (N n^3) so that long systems start
// first.
#include "../../include/socu_hip.h"
#include "nd_common.h"
#include "nd_internal.h"
#include "nd_tile.h"

namespace socu {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, int NT>
struct GroupScratch {
  T A[NT * (NT + 1)];
  T B[NT * (NT + 1)];
  T col[NT];
  T dinv[NT];
};

// Load an n x n row-major block into an NT x (NT+1) tile; entries outside n
// become the identity (diag) or zero.
template <typename T, int NT>
__device__ __forceinline__ void grp_load(T* tile, const T* __restrict__ src, int n, int gl, bool ident) {
  constexpr int TP = NT + 1;
  for (int idx = gl; idx < NT * NT; idx += 2 * NT) {
    const int r = idx / NT, c = idx - r * NT;
    tile[r * TP + c] = (r < n && c < n) ? src[(size_t)r * n + c] : ((ident && r == c) ? T(1) : T(0));
  }
}

template <typename T, int NT>
__device__ __forceinline__ void grp_load_rhs(T* tile, const T* __restrict__ src, int n, int m, int gl) {
  constexpr int TP = NT + 1;
  for (int idx = gl; idx < NT * m; idx += 2 * NT) {
    const int r = idx / m, c = idx - r * m;
    tile[r * TP + c] = (r < n) ? src[(size_t)r * m + c] : T(0);
  }
}

template <typename T, int NT, int W>
__device__ __forceinline__ void batched_system(const BatchProblem P, int op, GroupScratch<T, NT>* scr) {
  constexpr int G = 64 / (2 * NT);
  constexpr int TP = NT + 1;
  const int wave = wave_id();
  const int lane = lane_id();
  const int g = lane / (2 * NT);
  const int gl = lane % (2 * NT);
  const bool rowl = gl < NT;
  const int r = gl % NT;
  GroupScratch<T, NT>& my = scr[wave * G + g];
  const int N = P.N, n = P.n, m = P.m;
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  T* L = static_cast<T*>(P.L);
  T* E = static_cast<T*>(P.E);
  T* x = static_cast<T*>(P.x);
  const bool do_factor = op == SOCU_OP_FACTOR || op == SOCU_OP_FACTOR_AND_SOLVE;
  const bool do_fwd = op == SOCU_OP_FACTOR_AND_SOLVE || op == SOCU_OP_SOLVE || op == SOCU_OP_FORWARD;
  const bool do_bwd = op == SOCU_OP_FACTOR_AND_SOLVE || op == SOCU_OP_SOLVE || op == SOCU_OP_BACKWARD;
  const int Lv = recursive_iterations(N);
  const int slots = W * G;  // concurrent tasks per workgroup
  constexpr int kStride = (int)(sizeof(GroupScratch<T, NT>) / sizeof(T));
  static_assert(sizeof(GroupScratch<T, NT>) % sizeof(T) == 0, "scratch stride");
  GroupScratch<T, NT>& w0 = scr[wave * G];

  if (do_factor || do_fwd) {
    int off = 0;
    for (int l = 0; l < Lv; ++l) {
      const int s = 1 << l;
      const int dim = (N + s) / (2 * s);
      const int next_off = off + (N >> l) - 1;
      // ---- node tasks ----
      for (int base = 0; base < dim; base += slots) {
        const int t = base + wave * G + g;
        const bool real = t < dim;
        const int i = s - 1 + 2 * s * t;
        T* Li = L + (size_t)i * nn;
        if (do_factor) {
          if (real) grp_load<T, NT>(my.A, Li, n, gl, true);
          else
            for (int idx = gl; idx < NT * NT; idx += 2 * NT) my.A[(idx / NT) * TP + idx % NT] = (idx / NT == idx % NT);
          wave_sync();
          tile_chol_inv_grouped<T, NT>(w0.A, w0.col, w0.dinv, kStride, TP);
          wave_sync();
          if (real)
            for (int idx = gl; idx < n * n; idx += 2 * NT) Li[idx] = my.A[(idx / n) * TP + idx % n];
        } else if (real) {
          // factor already in L: bring L_i (lower + stored inverse) into the tile
          grp_load<T, NT>(my.A, Li, n, gl, true);
          if (rowl) my.dinv[r] = (r < n) ? T(1) / Li[(size_t)r * n + r] : T(1);
          wave_sync();
        }
        if (real) {
          // Linv[c][k] = A[k][c] (k < c), dinv[c] (k == c)
          if (do_factor && i + s < N && rowl && r < n) {  // R <- R Linv^T, lane r holds row r of R
            T* R = E + (size_t)(off + i / s) * nn;
            T row[NT];
#pragma unroll
            for (int k = 0; k < NT; ++k) row[k] = (k < n) ? R[(size_t)r * n + k] : T(0);
#pragma unroll
            for (int c = 0; c < NT; ++c) {
              T acc = row[c] * my.dinv[c];
#pragma unroll
              for (int k = 0; k < c; ++k) acc += row[k] * my.A[k * TP + c];
              if (c < n) R[(size_t)r * n + c] = acc;
            }
          }
          if (do_fwd) {  // y_i <- Linv y_i
            grp_load_rhs<T, NT>(my.B, x + (size_t)i * nm, n, m, gl);
            wave_sync();
            if (rowl && r < n)
              for (int c = 0; c < m; ++c) {
                T acc = my.dinv[r] * my.B[r * TP + c];
                for (int k = 0; k < r; ++k) acc += my.A[k * TP + r] * my.B[k * TP + c];
                x[(size_t)i * nm + (size_t)r * m + c] = acc;
              }
            wave_sync();
          }
          if (do_factor && i >= s) {  // Lc <- Linv Lc
            T* Lc = E + (size_t)(off + (i - s) / s) * nn;
            grp_load<T, NT>(my.B, Lc, n, gl, false);
            wave_sync();
            if (rowl && r < n)
              for (int c = 0; c < n; ++c) {
                T acc = my.dinv[r] * my.B[r * TP + c];
                for (int k = 0; k < r; ++k) acc += my.A[k * TP + r] * my.B[k * TP + c];
                Lc[(size_t)r * n + c] = acc;
              }
            wave_sync();
          }
        }
      }
      __syncthreads();
      // ---- receiving-block tasks: targets j = 2s(u+1)-1, then fill-ins ----
      const int nT = N / (2 * s);
      int nF = 0;
      if (do_factor)
        for (int tt = 1; tt < dim; ++tt)
          if (s - 1 + 2 * s * tt + s < N) ++nF;
      const int ntask = nT + nF;
      for (int base = 0; base < ntask; base += slots) {
        const int u = base + wave * G + g;
        if (u < nT) {
          const int j = 2 * s * (u + 1) - 1;
          const T* R = E + (size_t)(off + (j - s) / s) * nn;  // block (j, j-s)
          const bool right = j + s < N;
          const T* Lc = right ? E + (size_t)(off + j / s) * nn : nullptr;  // block (j+s, j)
          grp_load<T, NT>(my.A, R, n, gl, false);
          if (right) grp_load<T, NT>(my.B, Lc, n, gl, false);
          wave_sync();
          if (rowl && r < n) {
            if (do_factor) {
              T* Lj = L + (size_t)j * nn;
              for (int c = 0; c <= r; ++c) {
                T acc = T(0);
                for (int k = 0; k < n; ++k) acc += my.A[r * TP + k] * my.A[c * TP + k];
                if (right)
                  for (int k = 0; k < n; ++k) acc += my.B[k * TP + r] * my.B[k * TP + c];
                Lj[(size_t)r * n + c] -= acc;
              }
            }
            if (do_fwd) {
              const T* yl = x + (size_t)(j - s) * nm;
              const T* yr = right ? x + (size_t)(j + s) * nm : nullptr;
              for (int c = 0; c < m; ++c) {
                T acc = T(0);
                for (int k = 0; k < n; ++k) acc += my.A[r * TP + k] * yl[(size_t)k * m + c];
                if (right)
                  for (int k = 0; k < n; ++k) acc += my.B[k * TP + r] * yr[(size_t)k * m + c];
                x[(size_t)j * nm + (size_t)r * m + c] -= acc;
              }
            }
          }
          wave_sync();
        } else if (u < ntask) {
          const int tt = u - nT + 1;
          const int i = s - 1 + 2 * s * tt;
          const T* R = E + (size_t)(off + i / s) * nn;
          const T* Lc = E + (size_t)(off + (i - s) / s) * nn;
          T* F = E + (size_t)(next_off + (i - s) / (2 * s)) * nn;
          grp_load<T, NT>(my.A, R, n, gl, false);
          grp_load<T, NT>(my.B, Lc, n, gl, false);
          wave_sync();
          if (rowl && r < n)
            for (int c = 0; c < n; ++c) {
              T acc = T(0);
              for (int k = 0; k < n; ++k) acc += my.A[r * TP + k] * my.B[k * TP + c];
              F[(size_t)r * n + c] -= acc;
            }
          wave_sync();
        }
      }
      __syncthreads();
      off = next_off;
    }
  }

  if (do_bwd) {
    for (int l = Lv - 1; l >= 0; --l) {
      const int s = 1 << l;
      const int dim = (N + s) / (2 * s);
      int off = 0;
      for (int q = 0; q < l; ++q) off += (N >> q) - 1;
      for (int base = 0; base < dim; base += slots) {
        const int t = base + wave * G + g;
        if (t < dim) {
          const int i = s - 1 + 2 * s * t;
          const T* Li = L + (size_t)i * nn;
          const bool right = i + s < N, left = i >= s;
          // A <- R (block (i+s, i)) used as R^T; B <- updated x_i
          if (right) grp_load<T, NT>(my.A, E + (size_t)(off + i / s) * nn, n, gl, false);
          grp_load_rhs<T, NT>(my.B, x + (size_t)i * nm, n, m, gl);
          wave_sync();
          // x_i -= R^T x_{i+s} + Lc x_{i-s}: each row lane updates its own row of B
          if (rowl && r < n) {
            const T* Lc = left ? E + (size_t)(off + (i - s) / s) * nn : nullptr;  // block (i, i-s)
            const T* xr = right ? x + (size_t)(i + s) * nm : nullptr;
            const T* xl = left ? x + (size_t)(i - s) * nm : nullptr;
            for (int c = 0; c < m; ++c) {
              T acc = my.B[r * TP + c];
              if (right)
                for (int k = 0; k < n; ++k) acc -= my.A[k * TP + r] * xr[(size_t)k * m + c];
              if (left)
                for (int k = 0; k < n; ++k) acc -= Lc[(size_t)r * n + k] * xl[(size_t)k * m + c];
              my.B[r * TP + c] = acc;
            }
          }
          wave_sync();
          // x_i <- Linv^T x_i:  out[r] = x[r] / L[r][r] + sum_{k>r} U[r][k] x[k]
          if (rowl && r < n) {
            const T dg = Li[(size_t)r * n + r];
            for (int c = 0; c < m; ++c) {
              T acc = my.B[r * TP + c] / dg;
              for (int k = r + 1; k < n; ++k) acc += Li[(size_t)r * n + k] * my.B[k * TP + c];
              x[(size_t)i * nm + (size_t)r * m + c] = acc;
            }
          }
          wave_sync();
        }
      }
      __syncthreads();
    }
  }
}

// One workgroup of W wavefronts per system of tile class NT; `order` lists
// the systems of the class by decreasing cost.
template <typename T, int NT, int W>
__global__ __launch_bounds__(64 * W) void k_batched(const BatchProblem* __restrict__ probs,
                                                   const int* __restrict__ order, const uint8_t* __restrict__ mask,
                                                   int op) {
  const int p = order[blockIdx.x];
  if (masked_out(mask, p)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  batched_system<T, NT, W>(probs[p], op, reinterpret_cast<GroupScratch<T, NT>*>(smem_raw));
}

template <typename T, int NT, int W>
constexpr size_t batched_lds() {
  return sizeof(GroupScratch<T, NT>) * W * (64 / (2 * NT));
}

template <typename T, int NT, int W>
static hipError_t set_attr() {
  return hipFuncSetAttribute((const void*)k_batched<T, NT, W>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)batched_lds<T, NT, W>());
}

template <typename T, int NT>
static hipError_t set_attrs_nt() {
  hipError_t e = set_attr<T, NT, 1>();
  if (e == hipSuccess) e = set_attr<T, NT, 2>();
  if (e == hipSuccess) e = set_attr<T, NT, 4>();
  return e;
}

hipError_t init_batched_kernels() {
  hipError_t e = set_attrs_nt<double, 8>();
  if (e == hipSuccess) e = set_attrs_nt<double, 16>();
  if (e == hipSuccess) e = set_attrs_nt<double, 32>();
  if (e == hipSuccess) e = set_attrs_nt<float, 8>();
  if (e == hipSuccess) e = set_attrs_nt<float, 16>();
  if (e == hipSuccess) e = set_attrs_nt<float, 32>();
  return e;
}

int batched_class(int n) { return n <= 8 ? 0 : (n <= 16 ? 1 : 2); }

template <typename T, int NT, int W>
static void launch_w(const BatchProblem* dprobs, const int* dorder, int count, const uint8_t* mask, int op,
                     hipStream_t st) {
  const size_t lds = batched_lds<T, NT, W>();
  hipLaunchKernelGGL((k_batched<T, NT, W>), dim3(count), dim3(64 * W), lds, st, dprobs, dorder, mask, op);
}

template <typename T, int NT>
static void launch_nt(const BatchProblem* dprobs, const int* dorder, int count, int waves, const uint8_t* mask,
                      int op, hipStream_t st) {
  if (waves == 1) launch_w<T, NT, 1>(dprobs, dorder, count, mask, op, st);
  else if (waves == 2) launch_w<T, NT, 2>(dprobs, dorder, count, mask, op, st);
  else launch_w<T, NT, 4>(dprobs, dorder, count, mask, op, st);
}

hipError_t batched_launch_class(const BatchProblem* dprobs, const int* dorder, int cls, int count, int waves,
                                const uint8_t* mask, int op, int dtype, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (dtype == SOCU_DTYPE_F64) {
    if (cls == 0) launch_nt<double, 8>(dprobs, dorder, count, waves, mask, op, st);
    else if (cls == 1) launch_nt<double, 16>(dprobs, dorder, count, waves, mask, op, st);
    else launch_nt<double, 32>(dprobs, dorder, count, waves, mask, op, st);
  } else {
    if (cls == 0) launch_nt<float, 8>(dprobs, dorder, count, waves, mask, op, st);
    else if (cls == 1) launch_nt<float, 16>(dprobs, dorder, count, waves, mask, op, st);
    else launch_nt<float, 32>(dprobs, dorder, count, waves, mask, op, st);
  }
  return hipGetLastError();
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
