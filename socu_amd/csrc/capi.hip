// C-ABI of libsocu_hip.so: argument validation, the ND level loops, plans and
// hipGraph capture.  The level loops here replace the reference's host
// callbacks (block_tridiag_solver.py:792-813, :864-918, :992-1028,
// :1124-1228, :1294-1310, :1339-1369, :1435-1451, :1480-1512, :1586-1616,
// :1666-1722); no Python runs per level.
#include <stdarg.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <stdlib.h>
#include <string>
#include <vector>

#include "../../include/socu_hip.h"
#include "nd_common.h"
#include "nd_internal.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(SOCU_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int check_dims(const void* L, const void* E, const void* x, bool need_x, int B, int N, int n, int m, int dtype) {
  if (dtype != SOCU_DTYPE_F32 && dtype != SOCU_DTYPE_F64) return fail(SOCU_EINVAL, "dtype must be f32 or f64");
  if (B < 0) return fail(SOCU_EINVAL, "batch dimension must be >= 0 (got %d)", B);
  if (N < 1) return fail(SOCU_EINVAL, "horizon N must be >= 1 (got %d)", N);
  if (n < 1) return fail(SOCU_EINVAL, "block size n must be >= 1 (got %d)", n);
  if (n > socu::kMaxNLarge)
    return fail(SOCU_EUNSUPPORTED, "block size n=%d exceeds the supported maximum %d", n, socu::kMaxNLarge);
  if (need_x && m < 1) return fail(SOCU_EINVAL, "number of right-hand sides m must be >= 1 (got %d)", m);
  if (B > 0) {
    if (L == nullptr) return fail(SOCU_EINVAL, "L is NULL");
    if (E == nullptr && N > 1) return fail(SOCU_EINVAL, "E is NULL");
    if (need_x && x == nullptr) return fail(SOCU_EINVAL, "x is NULL");
  }
  if (B > 65535) return fail(SOCU_EUNSUPPORTED, "batch dimension %d exceeds 65535", B);
  // per-system offsets inside the kernels are 32-bit
  if ((size_t)N * n * n > 0x7fffffffu)
    return fail(SOCU_EUNSUPPORTED, "N*n*n = %zu exceeds 2^31 - 1 elements per system", (size_t)N * n * n);
  return SOCU_OK;
}

template <typename T>
int run_factor(T* L, T* E, const uint8_t* mask, int B, int N, int n, hipStream_t st, const T* Lsrc = nullptr,
               const T* Esrc = nullptr) {
  const int Lv = socu::recursive_iterations(N);
  for (int l = 0; l < Lv; ++l) {
    hipError_t e = socu::factor_level_blocked<T>(L, E, mask, B, N, n, l, false, st, nullptr, nullptr, nullptr,
                                                 l == 0 ? Lsrc : nullptr, l == 0 ? Esrc : nullptr);
    if (e != hipSuccess) return hip_fail(e, "factor level launch");
  }
  return SOCU_OK;
}

template <typename T>
int run_forward(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, hipStream_t st) {
  const int Lv = socu::recursive_iterations(N);
  for (int step = -1; step < Lv - 1; ++step) {
    hipError_t e = socu::forward_step_blocked<T>(L, E, x, mask, B, N, n, m, step, st);
    if (e != hipSuccess) return hip_fail(e, "forward step launch");
  }
  return SOCU_OK;
}

template <typename T>
int run_backward(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, hipStream_t st) {
  const int Lv = socu::recursive_iterations(N);
  for (int l = Lv - 1; l >= 0; --l) {
    hipError_t e = socu::backward_level_blocked<T>(L, E, x, mask, B, N, n, m, l, st);
    if (e != hipSuccess) return hip_fail(e, "backward level launch");
  }
  return SOCU_OK;
}

// Workspace of the persistent sweeps (owned by a plan); both NULL selects
// the level-synchronous sweep kernels.
struct SweepWs {
  int* ws = nullptr;
  void* slots = nullptr;
  int* hstatus = nullptr;      // pinned host word an expired wait raises
  int* hstatus_dev = nullptr;  // its device address
};

template <typename T>
int run_sweeps(const T* L, const T* E, T* x, const uint8_t* mask, int B, int N, int n, int m, bool fwd, bool bwd,
               const SweepWs& sw, hipStream_t st) {
  if (sw.ws != nullptr && m == 1) {
    hipError_t e = socu::sweep_persistent<T>(L, E, x, mask, B, N, n, fwd, bwd, sw.ws, static_cast<T*>(sw.slots),
                                             sw.hstatus_dev, st);
    if (e != hipSuccess) return hip_fail(e, "persistent sweep launch");
    return SOCU_OK;
  }
  int rc = SOCU_OK;
  if (fwd) rc = run_forward<T>(L, E, x, mask, B, N, n, m, st);
  if (rc == SOCU_OK && bwd) rc = run_backward<T>(L, E, x, mask, B, N, n, m, st);
  return rc;
}

template <typename T>
int run_op(int op, void* L, void* E, void* x, const uint8_t* mask, int B, int N, int n, int m, const SweepWs& sw,
           const socu::SweepOverlap* ov, void* inv, hipStream_t st, const void* Lsrc = nullptr,
           const void* Esrc = nullptr) {
  T* Lt = static_cast<T*>(L);
  T* Et = static_cast<T*>(E);
  T* xt = static_cast<T*>(x);
  const T* Ls = static_cast<const T*>(Lsrc);
  const T* Es = static_cast<const T*>(Esrc);
  int rc = SOCU_OK;
  if (Ls != nullptr && n > socu::kMaxN) {
    // the tiled kernels of n > 128 factor in place only: copy, then factor
    const size_t nn = (size_t)n * n, es = sizeof(T);
    hipError_t e = hipMemcpyAsync(Lt, Ls, (size_t)B * N * nn * es, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess && N > 1) {
      const size_t pitch = (size_t)socu::off_diag_storage_len(N) * nn * es;
      e = hipMemcpy2DAsync(Et, pitch, Es, pitch, (size_t)(N - 1) * nn * es, B, hipMemcpyDeviceToDevice, st);
    }
    if (e != hipSuccess) return hip_fail(e, "input copy of the out-of-place factor");
    Ls = nullptr;
    Es = nullptr;
  }
  if ((op == SOCU_OP_FORWARD || op == SOCU_OP_SOLVE || op == SOCU_OP_BACKWARD) && inv != nullptr) {
    // The sweeps multiply by the inverses of the diagonal factor blocks;
    // they are formed here from tril(L), so the factor may come from
    // anywhere (only tril(L) is read, as in the reference), and the solve
    // reads them from the workspace in place of L.  (Without a workspace
    // the caller vouches that L holds them already, as this library's
    // factor leaves them.)
    T* W = static_cast<T*>(inv);
    const hipError_t e = socu::trtri_blocked<T>(Lt, W, mask, B, N, n, st);
    if (e != hipSuccess) return hip_fail(e, "diagonal-block inverse launch");
    Lt = W;
  }
  switch (op) {
    case SOCU_OP_FACTOR:
      return run_factor<T>(Lt, Et, mask, B, N, n, st, Ls, Es);
    case SOCU_OP_FORWARD:
      return run_sweeps<T>(Lt, Et, xt, mask, B, N, n, m, true, false, sw, st);
    case SOCU_OP_BACKWARD:
      return run_sweeps<T>(Lt, Et, xt, mask, B, N, n, m, false, true, sw, st);
    case SOCU_OP_SOLVE:
      return run_sweeps<T>(Lt, Et, xt, mask, B, N, n, m, true, true, sw, st);
    case SOCU_OP_FACTOR_AND_SOLVE:
      if (socu::fused_forward_applies<T>(n, 0)) {  // forward sweep inside the factor's kernels
        const hipError_t e = socu::factor_and_forward<T>(Lt, Et, xt, mask, B, N, n, m, 0, st, Ls, Es);
        if (e != hipSuccess) return hip_fail(e, "factor_and_solve launch");
        return run_sweeps<T>(Lt, Et, xt, mask, B, N, n, m, false, true, sw, st);
      }
      if (ov != nullptr) {  // forward sweep alongside the factor, then the backward sweep
        const hipError_t e =
            socu::factor_with_forward<T>(Lt, Et, xt, mask, mask, B, N, n, m, false, *ov, st, nullptr, Ls, Es);
        if (e != hipSuccess) return hip_fail(e, "factor_and_solve launch");
        return run_sweeps<T>(Lt, Et, xt, mask, B, N, n, m, false, true, sw, st);
      }
      rc = run_factor<T>(Lt, Et, mask, B, N, n, st, Ls, Es);
      if (rc) return rc;
      return run_sweeps<T>(Lt, Et, xt, mask, B, N, n, m, true, true, sw, st);
    default:
      return fail(SOCU_EINVAL, "unknown operation %d", op);
  }
}

// Runs the per-process kernel set-up once; returns SOCU_OK or an error code.
int ensure_init() {
  static int state = 0;  // 0 = not yet, 1 = done
  if (state == 1) return SOCU_OK;
  hipError_t e = socu::init_blocked_kernels();
  if (e == hipSuccess) e = socu::init_refine_kernels();
  if (e == hipSuccess) e = socu::init_batched_kernels();
  if (e == hipSuccess) e = socu::init_sweep_kernels();
  if (e == hipSuccess) e = socu::init_large_kernels();
  if (e != hipSuccess) return hip_fail(e, "kernel set-up (hipFuncSetAttribute)");
  state = 1;
  return SOCU_OK;
}

// Bytes of the diagonal-block inverse workspace of a solve, forward or
// backward substitution: one n x n block per diagonal block.
size_t inverse_ws_bytes(int B, int N, int n, int dtype) {
  const size_t es = dtype == SOCU_DTYPE_F64 ? sizeof(double) : sizeof(float);
  return (size_t)B * N * n * n * es;
}

bool op_needs_inverse(int op) { return op == SOCU_OP_FORWARD || op == SOCU_OP_SOLVE || op == SOCU_OP_BACKWARD; }

int dispatch(int op, void* L, void* E, void* x, const unsigned char* mask, int B, int N, int n, int m, int dtype,
             void* stream, void* inv, const SweepWs& sw = SweepWs(), const socu::SweepOverlap* ov = nullptr,
             bool stored_inverse = false, const void* Lsrc = nullptr, const void* Esrc = nullptr) {
  const bool need_x = op != SOCU_OP_FACTOR;
  int rc = check_dims(L, E, x, need_x, B, N, n, m, dtype);
  if (rc) return rc;
  if (B == 0) return SOCU_OK;
  if (op_needs_inverse(op) && inv == nullptr && !stored_inverse)
    return fail(SOCU_EINVAL, "work is NULL: substitutions need socu_solve_workspace_size() bytes of device workspace");
  rc = ensure_init();
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == SOCU_DTYPE_F64) return run_op<double>(op, L, E, x, mask, B, N, n, m, sw, ov, inv, st, Lsrc, Esrc);
  return run_op<float>(op, L, E, x, mask, B, N, n, m, sw, ov, inv, st, Lsrc, Esrc);
}

// Per-thread launch timer used by socu_profile_op.
struct TimerState {
  bool active = false;
  std::vector<int> cls;
  std::vector<hipEvent_t> ev;  // begin/end pairs
  int open = -1;
};
thread_local TimerState g_timer;

}  // namespace

namespace socu {

void kt_begin(int cls, hipStream_t st) {
  if (!g_timer.active) return;
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
  (void)hipEventRecord(a, st);
  g_timer.cls.push_back(cls);
  g_timer.ev.push_back(a);
  g_timer.ev.push_back(b);
}

bool kt_kernel_events(int cls, hipEvent_t* start, hipEvent_t* stop) {
  if (!g_timer.active) return false;
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess) return false;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return false;
  }
  g_timer.cls.push_back(cls);
  g_timer.ev.push_back(a);
  g_timer.ev.push_back(b);
  *start = a;
  *stop = b;
  return true;
}

void kt_end(int cls, hipStream_t st) {
  (void)cls;
  if (!g_timer.active || g_timer.ev.empty()) return;
  (void)hipEventRecord(g_timer.ev.back(), st);
}

}  // namespace socu

struct socu_plan {
  int op;
  void* L;
  void* E;
  void* x;
  const unsigned char* mask;
  int B, N, n, m, dtype;
  hipGraph_t graph;
  hipGraphExec_t exec;
  SweepWs sweep;  // persistent-sweep workspace (m == 1 solves), else NULL
  socu::SweepOverlap overlap;  // factor_and_solve: forward sweep alongside the factor
  bool has_overlap;
  void* inv;  // solve / forward / backward: diagonal-block inverses formed from tril(L)
};

namespace {

void free_sweep_ws(SweepWs& sw) {
  if (sw.ws) (void)hipFree(sw.ws);
  if (sw.slots) (void)hipFree(sw.slots);
  if (sw.hstatus) (void)hipHostFree(sw.hstatus);
  sw = SweepWs();
}

// Allocates the persistent-sweep workspace of a plan whose operation
// includes a substitution with one right-hand side.
int alloc_sweep_ws(SweepWs& sw, int op, int B, int N, int n, int m, int dtype) {
  const size_t esz = dtype == SOCU_DTYPE_F64 ? sizeof(double) : sizeof(float);
  if (op == SOCU_OP_FACTOR || m != 1 || B <= 0 || !socu::sweep_persistent_enabled(n, (int)esz)) return SOCU_OK;
  hipError_t e = hipMalloc(&sw.ws, sizeof(int) * socu::sweep_ws_ints(B, N));
  if (e == hipSuccess) e = hipMemset(sw.ws, 0, sizeof(int) * socu::sweep_ws_ints(B, N));  // status reads 0 before a run
  if (e == hipSuccess) e = hipMalloc(&sw.slots, esz * socu::sweep_slot_elems(B, N, n));
  if (e == hipSuccess) e = hipHostMalloc(&sw.hstatus, sizeof(int), hipHostMallocMapped);
  if (e == hipSuccess) {
    *sw.hstatus = 0;
    e = hipHostGetDevicePointer((void**)&sw.hstatus_dev, sw.hstatus, 0);
  }
  if (e != hipSuccess) {
    free_sweep_ws(sw);
    return hip_fail(e, "hipMalloc (sweep workspace)");
  }
  return SOCU_OK;
}

int sweep_status(int* ws, int* hstatus, int* persistent, int* expired, void* stream) {
  if (persistent == nullptr || expired == nullptr) return fail(SOCU_EINVAL, "output pointers are NULL");
  *persistent = ws != nullptr ? 1 : 0;
  *expired = 0;
  if (ws == nullptr) return SOCU_OK;
  hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  int st = 0;
  if (e == hipSuccess) e = hipMemcpy(&st, ws + 1, sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "reading the sweep status");
  *expired = st != 0 ? 1 : 0;
  if (hstatus != nullptr) __atomic_store_n(hstatus, 0, __ATOMIC_SEQ_CST);  // reported here
  return SOCU_OK;
}

// Scratch of socu_cholesky_solve_noalloc: the diagonal-block inverses its
// sweeps read (socu_cholesky_solve's `work`).  Every GPU the library runs on
// holds one copy from module load until the process ends; calls on one GPU
// share it, so calls that overlap on one GPU race.  Host code never passes
// `scratch` itself (its host-side name is not a device address): it uses the
// per-device address that hipGetSymbolAddress returns.
constexpr size_t kNoAllocScratchBytes = SOCU_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on
// that device (with it current) and kept.
int noalloc_scratch(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(SOCU_EUNSUPPORTED, "device %d beyond %d devices", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_addr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress (no-allocation scratch)");
    if (p == nullptr) return fail(SOCU_EHIP, "hipGetSymbolAddress returned NULL for the no-allocation scratch");
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return SOCU_OK;
}

// An execution whose persistent sweep had a wait expire left NaN in its
// solution and raised the host-visible word; the next execution reports it
// (and clears the word) instead of running.
int check_expired(int* hstatus) {
  if (hstatus == nullptr || __atomic_load_n(hstatus, __ATOMIC_SEQ_CST) == 0) return SOCU_OK;
  __atomic_store_n(hstatus, 0, __ATOMIC_SEQ_CST);
  return fail(SOCU_EEXPIRED,
              "a dependency wait of the persistent substitution sweep expired in an earlier call of this launch; "
              "that call's solution holds NaN");
}

}  // namespace

extern "C" {

int socu_calculate_recursive_iterations(int N) {
  if (N < 1) return fail(SOCU_EINVAL, "horizon must be >= 1");
  return socu::recursive_iterations(N);
}

long long socu_calculate_off_diag_storage_len(int N) {
  if (N < 1) return fail(SOCU_EINVAL, "horizon must be >= 1");
  return socu::off_diag_storage_len(N);
}

int socu_max_block_size(void) { return socu::kMaxNLarge; }

int socu_set_sweep_spin_limit(int polls) {
  socu::set_sweep_spin_limit(polls);
  return SOCU_OK;
}

const char* socu_last_error(void) { return g_last_error.c_str(); }

size_t socu_solve_workspace_size(int B, int N, int n, int dtype) {
  if (B < 0 || N < 1 || n < 1 || (dtype != SOCU_DTYPE_F32 && dtype != SOCU_DTYPE_F64)) return 0;
  return inverse_ws_bytes(B, N, n, dtype);
}

int socu_cholesky_factor(void* L, void* E, const unsigned char* mask, int B, int N, int n, int dtype, void* stream) {
  return dispatch(SOCU_OP_FACTOR, L, E, nullptr, mask, B, N, n, 1, dtype, stream, nullptr);
}

int socu_cholesky_factor_f32(void* L, void* E, const unsigned char* mask, int B, int N, int n, int precision,
                             void* stream) {
  int rc = check_dims(L, E, nullptr, false, B, N, n, 1, SOCU_DTYPE_F32);
  if (rc) return rc;
  if (n > socu::kMaxN)
    return fail(SOCU_EUNSUPPORTED, "socu_cholesky_factor_f32: block size n=%d exceeds %d", n, socu::kMaxN);
  const int prec = precision == SOCU_PREC_FP32 ? socu::kPrecF64Pivots
                   : precision == SOCU_PREC_BF16 ? socu::kPrecF64Pivots | socu::kPrecBf16Schur
                   : precision == SOCU_PREC_FP32_F32PIV ? 0
                                                        : -1;
  if (prec < 0) return fail(SOCU_EINVAL, "unknown factor precision %d", precision);
  if (B == 0) return SOCU_OK;
  rc = ensure_init();
  if (rc) return rc;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const int Lv = socu::recursive_iterations(N);
  for (int l = 0; l < Lv; ++l) {
    const hipError_t e =
        socu::factor_level_blocked<float>(static_cast<float*>(L), static_cast<float*>(E), mask, B, N, n, l, prec, st);
    if (e != hipSuccess) return hip_fail(e, "factor level launch");
  }
  return SOCU_OK;
}

int socu_cholesky_forward_substitution(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N,
                                       int n, int m, int dtype, void* work, void* stream) {
  return dispatch(SOCU_OP_FORWARD, const_cast<void*>(L), const_cast<void*>(E), x, mask, B, N, n, m, dtype, stream,
                  work);
}

int socu_cholesky_backward_substitution(const void* L, const void* E, void* x, const unsigned char* mask, int B,
                                        int N, int n, int m, int dtype, void* work, void* stream) {
  return dispatch(SOCU_OP_BACKWARD, const_cast<void*>(L), const_cast<void*>(E), x, mask, B, N, n, m, dtype, stream,
                  work);
}

int socu_cholesky_solve(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N, int n, int m,
                        int dtype, void* work, void* stream) {
  return dispatch(SOCU_OP_SOLVE, const_cast<void*>(L), const_cast<void*>(E), x, mask, B, N, n, m, dtype, stream,
                  work);
}

int socu_cholesky_solve_with_inverses(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N,
                                      int n, int m, int dtype, void* stream) {
  return dispatch(SOCU_OP_SOLVE, const_cast<void*>(L), const_cast<void*>(E), x, mask, B, N, n, m, dtype, stream,
                  nullptr, SweepWs(), nullptr, true);
}

int socu_cholesky_solve_noalloc(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N, int n,
                                int m, int dtype, void* stream) {
  int rc = check_dims(L, E, x, true, B, N, n, m, dtype);
  if (rc) return rc;
  if (B == 0) return SOCU_OK;
  const size_t need = inverse_ws_bytes(B, N, n, dtype);
  if (need > kNoAllocScratchBytes)
    return fail(SOCU_EUNSUPPORTED,
                "socu_cholesky_solve_noalloc: B=%d N=%d n=%d needs %zu bytes of diagonal-block inverses, more than its "
                "%zu-byte scratch; use socu_cholesky_solve with a caller-provided workspace",
                B, N, n, need, kNoAllocScratchBytes);
  void* ws = nullptr;
  rc = noalloc_scratch(&ws);
  if (rc) return rc;
  return dispatch(SOCU_OP_SOLVE, const_cast<void*>(L), const_cast<void*>(E), x, mask, B, N, n, m, dtype, stream, ws);
}

int socu_cholesky_factor_and_solve(void* L, void* E, void* x, const unsigned char* mask, int B, int N, int n, int m,
                                   int dtype, void* stream) {
  return dispatch(SOCU_OP_FACTOR_AND_SOLVE, L, E, x, mask, B, N, n, m, dtype, stream, nullptr);
}

// the sources of an out-of-place factor: given, and disjoint from L / E
static int check_sources(const void* Ls, const void* Es, const void* L, const void* E, int B, int N, int n,
                         int dtype) {
  if (B == 0) return SOCU_OK;
  if (Ls == nullptr || (Es == nullptr && N > 1)) return fail(SOCU_EINVAL, "L_src or E_src is NULL");
  const size_t es = dtype == SOCU_DTYPE_F64 ? sizeof(double) : sizeof(float);
  const size_t nn = (size_t)n * n * es;
  auto overlap = [](const void* a, size_t na, const void* b, size_t nb) {
    const char* p = static_cast<const char*>(a);
    const char* q = static_cast<const char*>(b);
    return p < q + nb && q < p + na;
  };
  const size_t nL = (size_t)B * N * nn, nE = (size_t)B * socu::off_diag_storage_len(N) * nn;
  if (overlap(Ls, nL, L, nL) || (N > 1 && (overlap(Es, nE, E, nE) || overlap(Ls, nL, E, nE) || overlap(Es, nE, L, nL))))
    return fail(SOCU_EINVAL, "L_src / E_src overlap L / E");
  return SOCU_OK;
}

int socu_cholesky_factor_from(const void* L_src, const void* E_src, void* L, void* E, const unsigned char* mask, int B,
                              int N, int n, int dtype, void* stream) {
  int rc = check_dims(L, E, nullptr, false, B, N, n, 1, dtype);
  if (rc == SOCU_OK) rc = check_sources(L_src, E_src, L, E, B, N, n, dtype);
  if (rc) return rc;
  return dispatch(SOCU_OP_FACTOR, L, E, nullptr, mask, B, N, n, 1, dtype, stream, nullptr, SweepWs(), nullptr, false,
                  L_src, E_src);
}

int socu_cholesky_factor_and_solve_from(const void* L_src, const void* E_src, void* L, void* E, void* x,
                                        const unsigned char* mask, int B, int N, int n, int m, int dtype,
                                        void* stream) {
  int rc = check_dims(L, E, x, true, B, N, n, m, dtype);
  if (rc == SOCU_OK) rc = check_sources(L_src, E_src, L, E, B, N, n, dtype);
  if (rc) return rc;
  return dispatch(SOCU_OP_FACTOR_AND_SOLVE, L, E, x, mask, B, N, n, m, dtype, stream, nullptr, SweepWs(), nullptr,
                  false, L_src, E_src);
}

int socu_adjoint_scatter(const void* bbar, const void* x, void* dL, void* dE, const unsigned char* mask, int B, int N,
                         int n, int m, int dtype, void* stream) {
  if (bbar == nullptr || dL == nullptr || (dE == nullptr && N > 1))
    return fail(SOCU_EINVAL, "bbar, dL or dE is NULL");
  int rc = check_dims(dL, dE, x, true, B, N, n, m, dtype);
  if (rc) return rc;
  if (B == 0) return SOCU_OK;
  const hipError_t e = socu::adjoint_scatter(bbar, x, dL, dE, mask, B, N, n, m, dtype, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "adjoint scatter launch");
  return SOCU_OK;
}

int socu_factor_adjoint_scatter(const void* u, const void* v, const void* p, const void* q, void* dL, void* dE,
                                const unsigned char* mask, int B, int N, int n, int m, int dtype, void* stream) {
  if (u == nullptr || v == nullptr || dL == nullptr || (dE == nullptr && N > 1))
    return fail(SOCU_EINVAL, "u, v, dL or dE is NULL");
  if ((p == nullptr) != (q == nullptr)) return fail(SOCU_EINVAL, "p and q must both be given or both be NULL");
  int rc = check_dims(dL, dE, u, true, B, N, n, m, dtype);
  if (rc) return rc;
  if (B == 0) return SOCU_OK;
  const hipError_t e =
      socu::factor_adjoint_scatter(u, v, p, q, dL, dE, mask, B, N, n, m, dtype, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "factor adjoint scatter launch");
  return SOCU_OK;
}

int socu_profile_op(int op, void* L, void* E, void* x, const unsigned char* mask, int B, int N, int n, int m,
                    int dtype, void* work, void* stream, double* kernel_ms, int* kernel_launches) {
  if (kernel_ms == nullptr || kernel_launches == nullptr) return fail(SOCU_EINVAL, "output arrays are NULL");
  for (int c = 0; c < socu::kKcCount; ++c) {
    kernel_ms[c] = 0.0;
    kernel_launches[c] = 0;
  }
  g_timer = TimerState();
  g_timer.active = true;
  int rc = dispatch(op, L, E, x, mask, B, N, n, m, dtype, stream, work);
  g_timer.active = false;
  hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  for (size_t q = 0; q < g_timer.cls.size(); ++q) {
    float ms = 0.f;
    if (rc == SOCU_OK && e == hipSuccess &&
        hipEventElapsedTime(&ms, g_timer.ev[2 * q], g_timer.ev[2 * q + 1]) == hipSuccess) {
      kernel_ms[g_timer.cls[q]] += ms;
      kernel_launches[g_timer.cls[q]] += 1;
    }
  }
  for (hipEvent_t ev : g_timer.ev) (void)hipEventDestroy(ev);
  g_timer = TimerState();
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize (profile)");
  return SOCU_OK;
}

int socu_plan_create(socu_plan** out, int op, void* L, void* E, void* x, const unsigned char* mask, int B, int N,
                     int n, int m, int dtype, int use_graph) {
  if (out == nullptr) return fail(SOCU_EINVAL, "plan output pointer is NULL");
  *out = nullptr;
  if (op < SOCU_OP_FACTOR || op > SOCU_OP_BACKWARD) return fail(SOCU_EINVAL, "unknown operation %d", op);
  const bool need_x = op != SOCU_OP_FACTOR;
  int rc = check_dims(L, E, x, need_x, B, N, n, need_x ? m : 1, dtype);
  if (rc) return rc;
  if (B > 0) {
    rc = ensure_init();
    if (rc) return rc;
  }
  socu_plan* p = new socu_plan{op, L, E, x, mask, B, N, n, need_x ? m : 1, dtype, nullptr, nullptr, SweepWs(),
                               socu::SweepOverlap(), false, nullptr};
  rc = alloc_sweep_ws(p->sweep, op, B, N, n, p->m, dtype);
  if (rc == SOCU_OK && op_needs_inverse(op) && B > 0) {
    const hipError_t e = hipMalloc(&p->inv, inverse_ws_bytes(B, N, n, dtype));
    if (e != hipSuccess) {
      p->inv = nullptr;
      rc = hip_fail(e, "hipMalloc (diagonal-block inverse workspace)");
    }
  }
  // The forward sweep alongside the factor pays inside a graph only: issued
  // eagerly, its event records and waits cost more host time than the
  // overlap saves (n=128 N=512 FP64: 1.19 ms without, 1.24 ms with).
  if (rc == SOCU_OK && op == SOCU_OP_FACTOR_AND_SOLVE && B > 0 && use_graph) {
    const hipError_t e = socu::overlap_create(p->overlap);
    if (e != hipSuccess) rc = hip_fail(e, "stream / event creation (factor_and_solve)");
    p->has_overlap = e == hipSuccess;
  }
  if (rc) {
    socu_plan_destroy(p);
    return rc;
  }
  if (use_graph && B > 0) {
    hipStream_t cs;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) {
      socu_plan_destroy(p);  // workspace and streams; no graph yet
      return hip_fail(e, "hipStreamCreate (graph capture)");
    }
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(cs);
      socu_plan_destroy(p);  // workspace and streams; no graph yet
      return hip_fail(e, "hipStreamBeginCapture");
    }
    rc = dispatch(op, L, E, x, mask, B, N, n, p->m, dtype, cs, p->inv, p->sweep,
                  p->has_overlap ? &p->overlap : nullptr);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc != SOCU_OK || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      socu_plan_destroy(p);  // workspace and streams; no graph yet
      return rc != SOCU_OK ? rc : hip_fail(e2, "hipStreamEndCapture");
    }
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      socu_plan_destroy(p);  // workspace and streams; no graph yet
      return hip_fail(e, "hipGraphInstantiate");
    }
    p->graph = g;
    p->exec = ex;
  }
  *out = p;
  return SOCU_OK;
}

int socu_plan_execute(socu_plan* p, void* stream) {
  if (p == nullptr) return fail(SOCU_EINVAL, "plan is NULL");
  if (p->B == 0) return SOCU_OK;
  const int rc = check_expired(p->sweep.hstatus);
  if (rc) return rc;
  if (p->exec) {
    hipError_t e = hipGraphLaunch(p->exec, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    return SOCU_OK;
  }
  return dispatch(p->op, p->L, p->E, p->x, p->mask, p->B, p->N, p->n, p->m, p->dtype, stream, p->inv, p->sweep,
                  p->has_overlap ? &p->overlap : nullptr);
}

int socu_plan_sweep_status(socu_plan* p, int* persistent, int* expired, void* stream) {
  if (p == nullptr) return fail(SOCU_EINVAL, "plan is NULL");
  return sweep_status(p->sweep.ws, p->sweep.hstatus, persistent, expired, stream);
}

int socu_plan_destroy(socu_plan* p) {
  if (p == nullptr) return SOCU_OK;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  free_sweep_ws(p->sweep);
  if (p->inv) (void)hipFree(p->inv);
  if (p->has_overlap) socu::overlap_destroy(p->overlap);
  delete p;
  return SOCU_OK;
}

struct socu_mixed {
  socu::MixedPlan p;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

int socu_mixed_create(socu_mixed** out, const void* L, const void* E, void* x, const unsigned char* mask, int B, int N,
                      int n, int m, int precision, double tol, int max_iter, int use_graph) {
  if (out == nullptr) return fail(SOCU_EINVAL, "output pointer is NULL");
  *out = nullptr;
  int rc = check_dims(L, E, x, true, B, N, n, m, SOCU_DTYPE_F64);
  if (rc) return rc;
  if (precision != SOCU_PREC_FP32 && precision != SOCU_PREC_FP32_F32PIV && precision != SOCU_PREC_BF16)
    return fail(SOCU_EINVAL, "unknown factor precision %d", precision);
  if (!(tol > 0.0)) return fail(SOCU_EINVAL, "tolerance must be positive");
  if (max_iter < 1 || max_iter > 100) return fail(SOCU_EINVAL, "max_iter must be in [1, 100] (got %d)", max_iter);
  if (n > socu::kMaxN)
    return fail(SOCU_EUNSUPPORTED, "mixed-precision solve: block size n=%d exceeds %d", n, socu::kMaxN);
  if (!socu::mixed_fits(n, m))
    return fail(SOCU_EUNSUPPORTED, "n*m = %d too large for the residual kernel (max 4608)", n * m);
  if (B > 0) {
    rc = ensure_init();
    if (rc) return rc;
  }
  socu_mixed* h = new socu_mixed();
  h->p.L = static_cast<const double*>(L);
  h->p.E = static_cast<const double*>(E);
  h->p.x = static_cast<double*>(x);
  h->p.mask = mask;
  h->p.B = B;
  h->p.N = N;
  h->p.n = n;
  h->p.m = m;
  h->p.tol = tol;
  h->p.max_iter = max_iter;
  h->p.prec = precision == SOCU_PREC_FP32 ? socu::kPrecF64Pivots
               : precision == SOCU_PREC_BF16 ? socu::kPrecF64Pivots | socu::kPrecBf16Schur
                                             : 0;
  h->p.use_overlap = use_graph != 0;  // a second stream pays inside a graph only
  h->graph = nullptr;
  h->exec = nullptr;
  if (B == 0) {
    *out = h;
    return SOCU_OK;
  }
  hipError_t e = socu::mixed_alloc(h->p);
  if (e != hipSuccess) {
    socu::mixed_free(h->p);
    delete h;
    return hip_fail(e, "hipMalloc (mixed-precision workspace)");
  }
  if (use_graph) {
    hipStream_t cs;
    e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipError_t e1 = socu::mixed_enqueue(h->p, cs);
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(cs, &g);
      e = e1 != hipSuccess ? e1 : e2;
      if (e == hipSuccess) e = hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0);
      h->graph = g;
    }
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
      if (h->exec) (void)hipGraphExecDestroy(h->exec);
      if (h->graph) (void)hipGraphDestroy(h->graph);
      socu::mixed_free(h->p);
      delete h;
      return hip_fail(e, "graph capture (mixed-precision solve)");
    }
  }
  *out = h;
  return SOCU_OK;
}

int socu_mixed_execute(socu_mixed* h, void* stream) {
  if (h == nullptr) return fail(SOCU_EINVAL, "handle is NULL");
  if (h->p.B == 0) return SOCU_OK;
  const int rc = check_expired(h->p.sweep_hstatus);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = h->exec ? hipGraphLaunch(h->exec, st) : socu::mixed_enqueue(h->p, st);
  if (e != hipSuccess) return hip_fail(e, "mixed-precision solve launch");
  return SOCU_OK;
}

int socu_mixed_stats(socu_mixed* h, double* rel_resid, int* solves, void* stream) {
  if (h == nullptr || rel_resid == nullptr || solves == nullptr) return fail(SOCU_EINVAL, "NULL argument");
  const int B = h->p.B;
  if (B == 0) return SOCU_OK;
  std::vector<double> st(4 * (size_t)B);
  hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  if (e == hipSuccess) e = hipMemcpy(st.data(), h->p.state, sizeof(double) * 4 * B, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "reading refinement state");
  for (int b = 0; b < B; ++b) {
    const double bb = st[4 * b], rr = st[4 * b + 1];
    rel_resid[b] = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
    solves[b] = (int)st[4 * b + 3];
  }
  return SOCU_OK;
}

int socu_mixed_profile(socu_mixed* h, void* stream, double* kernel_ms, int* kernel_launches) {
  if (h == nullptr || kernel_ms == nullptr || kernel_launches == nullptr) return fail(SOCU_EINVAL, "NULL argument");
  for (int c = 0; c < socu::kKcCount; ++c) {
    kernel_ms[c] = 0.0;
    kernel_launches[c] = 0;
  }
  if (h->p.B == 0) return SOCU_OK;
  g_timer = TimerState();
  g_timer.active = true;
  hipError_t e = socu::mixed_enqueue(h->p, static_cast<hipStream_t>(stream));
  g_timer.active = false;
  hipError_t e2 = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  for (size_t q = 0; q < g_timer.cls.size(); ++q) {
    float ms = 0.f;
    if (e == hipSuccess && e2 == hipSuccess &&
        hipEventElapsedTime(&ms, g_timer.ev[2 * q], g_timer.ev[2 * q + 1]) == hipSuccess) {
      kernel_ms[g_timer.cls[q]] += ms;
      kernel_launches[g_timer.cls[q]] += 1;
    }
  }
  for (hipEvent_t ev : g_timer.ev) (void)hipEventDestroy(ev);
  g_timer = TimerState();
  if (e != hipSuccess) return hip_fail(e, "mixed-precision solve launch (profile)");
  if (e2 != hipSuccess) return hip_fail(e2, "hipStreamSynchronize (profile)");
  return SOCU_OK;
}

int socu_mixed_sweep_status(socu_mixed* h, int* persistent, int* expired, void* stream) {
  if (h == nullptr) return fail(SOCU_EINVAL, "handle is NULL");
  return sweep_status(h->p.sweep_ws, h->p.sweep_hstatus, persistent, expired, stream);
}

int socu_mixed_destroy(socu_mixed* h) {
  if (h == nullptr) return SOCU_OK;
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  socu::mixed_free(h->p);
  delete h;
  return SOCU_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched uneven systems.

namespace {

constexpr int kClasses = 3;  // tile classes NT = 8, 16, 32

// Process-wide tuning results: key -> chosen (tile class, wavefronts) coded
// as 8 * waves + class.  Optionally persisted in the file named by
// SOCU_TUNE_CACHE ("key code" per line).
std::mutex g_tune_mu;
std::map<std::string, int> g_tune;
bool g_tune_loaded = false;

void tune_cache_load() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* path = getenv("SOCU_TUNE_CACHE");
  if (path == nullptr || *path == 0) return;
  FILE* f = fopen(path, "r");
  if (f == nullptr) return;
  char key[256];
  int c = 0;
  while (fscanf(f, "%255s %d", key, &c) == 2)
    if (c >= 0 && (c & 7) < kClasses && ((c >> 3) == 1 || (c >> 3) == 2 || (c >> 3) == 4)) g_tune[key] = c;
  fclose(f);
}

void tune_cache_store(const std::string& key, int c) {
  g_tune[key] = c;
  const char* path = getenv("SOCU_TUNE_CACHE");
  if (path == nullptr || *path == 0) return;
  FILE* f = fopen(path, "a");
  if (f == nullptr) return;
  fprintf(f, "%s %d\n", key.c_str(), c);
  fclose(f);
}

}  // namespace

// Systems are grouped in size buckets b (max(n, m) <= 8, 16, 32); bucket b
// runs as one launch with tile class tile[b] >= b and workgroups of
// waves[b] wavefronts.
struct socu_batched {
  int op = 0, dtype = 0, count = 0;
  const unsigned char* mask = nullptr;
  socu::BatchProblem* dprobs = nullptr;  // device copy of the system table
  int* dorder = nullptr;                 // device: system indices grouped by bucket
  std::vector<int> order;                // host copy of dorder
  std::vector<int> bucket_of;            // per system
  std::vector<long long> cost;           // per system
  std::vector<int> mrhs;                 // per system
  int tile[kClasses] = {0, 1, 2};
  int waves[kClasses] = {1, 1, 1};
  int counts[kClasses] = {0, 0, 0};  // per bucket
  int starts[kClasses] = {0, 0, 0};
  int mmax[kClasses] = {1, 1, 1};
  long long bcost[kClasses] = {0, 0, 0};
  int launch_order[kClasses] = {2, 1, 0};  // buckets by decreasing total cost
  hipStream_t side[kClasses - 1] = {nullptr, nullptr};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kClasses - 1] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void batched_release(socu_batched* h) {
  if (h == nullptr) return;
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  for (int k = 0; k < kClasses - 1; ++k) {
    if (h->join[k]) (void)hipEventDestroy(h->join[k]);
    if (h->side[k]) (void)hipStreamDestroy(h->side[k]);
  }
  if (h->fork) (void)hipEventDestroy(h->fork);
  if (h->dprobs) (void)hipFree(h->dprobs);
  if (h->dorder) (void)hipFree(h->dorder);
  delete h;
}

// Groups the systems by bucket (each by decreasing cost: long systems start
// first) and uploads the order.
int batched_layout(socu_batched* h) {
  h->order.clear();
  for (int b = kClasses - 1; b >= 0; --b) {
    std::vector<int> o;
    for (int k = 0; k < h->count; ++k)
      if (h->bucket_of[k] == b) o.push_back(k);
    std::stable_sort(o.begin(), o.end(), [&](int x, int y) { return h->cost[x] > h->cost[y]; });
    h->starts[b] = (int)h->order.size();
    h->counts[b] = (int)o.size();
    h->mmax[b] = 1;
    h->bcost[b] = 0;
    for (int k : o) {
      h->mmax[b] = std::max(h->mmax[b], h->mrhs[k]);
      h->bcost[b] += h->cost[k];
    }
    h->order.insert(h->order.end(), o.begin(), o.end());
  }
  for (int b = 0; b < kClasses; ++b) h->launch_order[b] = b;
  std::sort(h->launch_order, h->launch_order + kClasses, [&](int x, int y) { return h->bcost[x] > h->bcost[y]; });
  const hipError_t e = hipMemcpy(h->dorder, h->order.data(), h->order.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "batched plan set-up (order)");
  return SOCU_OK;
}

// Launches every non-empty class: the costliest on st, the others on side
// streams forked from and joined back into st, so their workgroups fill the
// machine together.
int batched_enqueue(socu_batched* h, const unsigned char* mask, hipStream_t st,
                    const socu::BatchProblem* probs = nullptr) {
  if (probs == nullptr) probs = h->dprobs;
  int first = -1, nside = 0;
  hipError_t e = hipSuccess;
  for (int q = 0; q < kClasses && e == hipSuccess; ++q) {
    const int c = h->launch_order[q];
    if (h->counts[c] == 0) continue;
    if (first < 0) {
      first = c;
      continue;
    }
    if (nside == 0) e = hipEventRecord(h->fork, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(h->side[nside], h->fork, 0);
    if (e == hipSuccess)
      e = socu::batched_launch_class(probs, h->dorder + h->starts[c], h->tile[c], h->counts[c], h->mmax[c],
                                     h->waves[c], mask, h->op, h->dtype, h->side[nside]);
    if (e == hipSuccess) e = hipEventRecord(h->join[nside], h->side[nside]);
    ++nside;
  }
  if (e == hipSuccess && first >= 0)
    e = socu::batched_launch_class(probs, h->dorder + h->starts[first], h->tile[first], h->counts[first],
                                   h->mmax[first], h->waves[first], mask, h->op, h->dtype, st);
  for (int k = 0; k < nside && e == hipSuccess; ++k) e = hipStreamWaitEvent(st, h->join[k], 0);
  if (e != hipSuccess) return hip_fail(e, "batched launch");
  return SOCU_OK;
}

// For each size bucket, times every admissible (tile class >= the bucket's
// own, wavefronts 1, 2 or 4) on a device copy of the batch and keeps the
// fastest, near-ties going to the wider workgroup.  The caller's arrays are
// only read.  Needs batched_layout first.
int batched_tune(socu_batched* h, const std::vector<socu::BatchProblem>& probs) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_cache_load();
  std::string keys[kClasses];
  bool need = false;
  for (int b = 0; b < kClasses; ++b) {
    if (h->counts[b] == 0) continue;
    char buf[160];
    snprintf(buf, sizeof buf, "batched4:%s:op%d:bucket%d:count%d:cost%lld", h->dtype == SOCU_DTYPE_F64 ? "f64" : "f32",
             h->op, b, h->counts[b], h->bcost[b]);
    keys[b] = buf;
    auto it = g_tune.find(keys[b]);
    if (it != g_tune.end() && (it->second & 7) >= b &&
        socu::batched_fits(it->second & 7, it->second >> 3, h->mmax[b], h->dtype)) {
      h->tile[b] = it->second & 7;
      h->waves[b] = it->second >> 3;
    } else {
      need = true;
    }
  }
  if (!need) return SOCU_OK;
  const size_t es = h->dtype == SOCU_DTYPE_F64 ? sizeof(double) : sizeof(float);
  std::vector<socu::BatchProblem> copy(probs);
  size_t total = 0;
  std::vector<size_t> offs(probs.size());
  for (size_t k = 0; k < probs.size(); ++k) {
    const auto& P = probs[k];
    const size_t nn = (size_t)P.n * P.n;
    offs[k] = total;
    total += ((size_t)P.N * nn + (size_t)socu::off_diag_storage_len(P.N) * nn + (size_t)P.N * P.n * P.m) * es;
    total = (total + 255) & ~(size_t)255;
  }
  char* arena = nullptr;
  socu::BatchProblem* dcopy = nullptr;
  hipStream_t ts = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SOCU_OK;
  hipError_t e = hipMalloc(&arena, total);
  if (e == hipSuccess) {
    for (size_t k = 0; k < probs.size(); ++k) {
      const auto& P = probs[k];
      const size_t nn = (size_t)P.n * P.n;
      char* base = arena + offs[k];
      copy[k].L = base;
      copy[k].E = base + (size_t)P.N * nn * es;
      copy[k].x = (P.x != nullptr) ? base + ((size_t)P.N * nn + (size_t)socu::off_diag_storage_len(P.N) * nn) * es
                                   : nullptr;
    }
    e = hipMalloc(&dcopy, copy.size() * sizeof(socu::BatchProblem));
  }
  if (e == hipSuccess) e = hipMemcpy(dcopy, copy.data(), copy.size() * sizeof(socu::BatchProblem), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ts, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int b = 0; b < kClasses && e == hipSuccess; ++b) {
    if (h->counts[b] == 0 ||
        (g_tune.count(keys[b]) && (g_tune[keys[b]] & 7) >= b &&
         socu::batched_fits(g_tune[keys[b]] & 7, g_tune[keys[b]] >> 3, h->mmax[b], h->dtype)))
      continue;
    struct Cand {
      int c, w;
      float ms;
    };
    std::vector<Cand> cands;
    for (int c = b; c < kClasses && e == hipSuccess; ++c) {
      for (int w : {1, 2, 4}) {
        if (!socu::batched_fits(c, w, h->mmax[b], h->dtype)) continue;
        cands.push_back({c, w, 1e30f});
        for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
          e = socu::batched_copy(h->dprobs, dcopy, h->count, h->op, h->dtype, ts);
          if (e == hipSuccess) e = hipEventRecord(e0, ts);
          if (e == hipSuccess)
            e = socu::batched_launch_class(dcopy, h->dorder + h->starts[b], c, h->counts[b], h->mmax[b], w, nullptr,
                                           h->op, h->dtype, ts);
          if (e == hipSuccess) e = hipEventRecord(e1, ts);
          if (e == hipSuccess) e = hipEventSynchronize(e1);
          float ms = 0.f;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
          if (e == hipSuccess && ms < cands.back().ms) cands.back().ms = ms;
        }
      }
    }
    // Isolated times within kTuneTie of the fastest are ties, settled for
    // the most wavefronts per system: a wider workgroup runs more of a
    // system's tasks at once, which shortens the bucket's tail when the
    // buckets share the GPU in the real launch (measured on the 4096-system
    // MPC batch: the isolated picks of one and two wavefronts differ by 1-2%,
    // and two wavefronts for the large buckets make the whole batch 3-5%
    // faster).
    if (e == hipSuccess && !cands.empty()) {
      constexpr float kTuneTie = 0.02f;
      float tmin = 1e30f;
      for (const Cand& k : cands) tmin = std::min(tmin, k.ms);
      const Cand* pick = nullptr;
      for (const Cand& k : cands)
        if (k.ms <= tmin * (1.f + kTuneTie) && (pick == nullptr || k.w > pick->w || (k.w == pick->w && k.ms < pick->ms)))
          pick = &k;
      h->tile[b] = pick->c;
      h->waves[b] = pick->w;
    }
  }
  if (e == hipSuccess)
    for (int b = 0; b < kClasses; ++b)
      if (h->counts[b] > 0) tune_cache_store(keys[b], 8 * h->waves[b] + h->tile[b]);
  if (e != hipSuccess) rc = hip_fail(e, "batched autotune");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (ts) (void)hipStreamDestroy(ts);
  if (dcopy) (void)hipFree(dcopy);
  if (arena) (void)hipFree(arena);
  return rc;
}

}  // namespace

extern "C" {

int socu_batched_max_block_size(void) { return 32; }

int socu_batched_create(socu_batched** out, int op, const socu_batch_system* systems, int count,
                        const unsigned char* mask, int dtype, int tune, int use_graph) {
  if (out == nullptr) return fail(SOCU_EINVAL, "batched output pointer is NULL");
  *out = nullptr;
  if (op < SOCU_OP_FACTOR || op > SOCU_OP_BACKWARD) return fail(SOCU_EINVAL, "unknown operation %d", op);
  if (dtype != SOCU_DTYPE_F32 && dtype != SOCU_DTYPE_F64) return fail(SOCU_EINVAL, "unknown dtype code %d", dtype);
  if (count < 0) return fail(SOCU_EINVAL, "system count must be >= 0, got %d", count);
  if (count > 0 && systems == nullptr) return fail(SOCU_EINVAL, "system table is NULL");
  const bool need_x = op != SOCU_OP_FACTOR;
  std::vector<socu::BatchProblem> probs(count);
  socu_batched* h = new socu_batched();
  h->op = op;
  h->dtype = dtype;
  h->count = count;
  h->mask = mask;
  h->bucket_of.resize(count);
  h->cost.resize(count);
  h->mrhs.resize(count);
  for (int k = 0; k < count; ++k) {
    const socu_batch_system& S = systems[k];
    int rc = SOCU_OK;
    const int m = need_x ? S.m : 1;
    if (S.L == nullptr || (S.E == nullptr && S.N > 1)) rc = fail(SOCU_EINVAL, "system %d: L or E is NULL", k);
    else if (S.N < 1) rc = fail(SOCU_EINVAL, "system %d: horizon N must be >= 1, got %d", k, S.N);
    else if (S.n < 1) rc = fail(SOCU_EINVAL, "system %d: block size n must be >= 1, got %d", k, S.n);
    else if (S.reserved != 0) rc = fail(SOCU_EINVAL, "system %d: reserved field must be 0", k);
    else if (need_x && (S.x == nullptr || m < 1)) rc = fail(SOCU_EINVAL, "system %d: x is NULL or m < 1", k);
    else if (S.n > 32 || m > 32)
      rc = fail(SOCU_EUNSUPPORTED, "system %d: batched systems need n <= 32 and m <= 32 (got n=%d, m=%d)", k, S.n, m);
    if (rc) {
      delete h;
      return rc;
    }
    probs[k] = socu::BatchProblem{S.L, S.E, need_x ? S.x : nullptr, S.N, S.n, m, 0};
    h->cost[k] = (long long)S.N * S.n * S.n * (S.n + m);
    h->bucket_of[k] = socu::batched_class(std::max(S.n, m));
    h->mrhs[k] = m;
  }
  if (count == 0) {
    *out = h;
    return SOCU_OK;
  }
  int rc = ensure_init();
  if (rc) {
    batched_release(h);
    return rc;
  }
  hipError_t e = hipMalloc(&h->dprobs, count * sizeof(socu::BatchProblem));
  if (e == hipSuccess) e = hipMalloc(&h->dorder, count * sizeof(int));
  if (e == hipSuccess) e = hipMemcpy(h->dprobs, probs.data(), count * sizeof(socu::BatchProblem), hipMemcpyHostToDevice);
  for (int k = 0; k < kClasses - 1 && e == hipSuccess; ++k) {
    e = hipStreamCreateWithFlags(&h->side[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->join[k], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&h->fork, hipEventDisableTiming);
  if (e != hipSuccess) {
    batched_release(h);
    return hip_fail(e, "batched plan set-up");
  }
  rc = batched_layout(h);
  if (rc == SOCU_OK && !tune) {
    // SOCU_BATCHED_WAVES=1|2|4 sets the wavefronts per workgroup of an
    // untuned plan (every variant of the kernels can then be run on purpose),
    // narrowed where its LDS does not fit the bucket's right-hand sides.
    // Default 2: two of a system's tasks at a time, the tuner's usual pick
    // for the 16- and 32-wide buckets (4096-system MPC batch: 10.57 ms
    // against 10.7-10.8 ms with one wavefront)
    const char* w = getenv("SOCU_BATCHED_WAVES");
    const int wv = w ? atoi(w) : 2;
    if (wv == 1 || wv == 2 || wv == 4)
      for (int b = 0; b < kClasses; ++b) {
        int v = wv;
        while (v > 1 && !socu::batched_fits(h->tile[b], v, h->mmax[b], h->dtype)) v /= 2;
        h->waves[b] = v;
      }
  }
  if (rc == SOCU_OK && tune) rc = batched_tune(h, probs);
  if (rc) {
    batched_release(h);
    return rc;
  }
  if (use_graph) {
    hipStream_t cs;
    e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) {
      batched_release(h);
      return hip_fail(e, "hipStreamCreate (graph capture)");
    }
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(cs);
      batched_release(h);
      return hip_fail(e, "hipStreamBeginCapture");
    }
    rc = batched_enqueue(h, h->mask, cs);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc != SOCU_OK || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      batched_release(h);
      return rc != SOCU_OK ? rc : hip_fail(e2, "hipStreamEndCapture");
    }
    h->graph = g;
    e = hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      batched_release(h);
      return hip_fail(e, "hipGraphInstantiate");
    }
  }
  *out = h;
  return SOCU_OK;
}

int socu_batched_execute(socu_batched* h, void* stream) {
  if (h == nullptr) return fail(SOCU_EINVAL, "batched plan is NULL");
  if (h->count == 0) return SOCU_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (h->exec) {
    hipError_t e = hipGraphLaunch(h->exec, st);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    return SOCU_OK;
  }
  return batched_enqueue(h, h->mask, st);
}

int socu_batched_config(socu_batched* h, int* counts, int* tiles, int* waves) {
  if (h == nullptr || counts == nullptr || tiles == nullptr || waves == nullptr)
    return fail(SOCU_EINVAL, "NULL argument");
  for (int b = 0; b < kClasses; ++b) {
    counts[b] = h->counts[b];
    tiles[b] = 8 << h->tile[b];
    waves[b] = h->waves[b];
  }
  return SOCU_OK;
}

int socu_batched_destroy(socu_batched* h) {
  batched_release(h);
  return SOCU_OK;
}

}  // extern "C"
