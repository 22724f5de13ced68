// Allocation-free variant of the matrix-application entry point.
//
// Callers that may not allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) stage x and y in a
// static device array of this file instead of in per-call allocations.  The
// array is a host-side stand-in in host code: its device address is looked up
// once per device with hipGetSymbolAddress when a solver is created on that
// device, and only that address is passed to kernels and copies.
//
// Every call on one GPU shares that GPU's array and leaves the last call's
// contents in it; calls that overlap on one GPU (two host threads, or two
// solvers on different streams) race on it.
#include <mutex>

#include "ops.hpp"
#include "solver.hpp"

namespace cfsa {

// 64 MiB per device: x and y of one call together (8 Mi doubles).
constexpr size_t kScratchDoubles = (size_t)8 << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Device address of the scratch array on the current device, looked up on the
// first call per device (with that device current) and kept for the process.
double *noalloc_scratch_current_device() {
  int dev = 0;
  CFSA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    CFSA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw std::runtime_error("hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  return g_scratch_ptr[dev];
}

// y = A x as apply_matrix, with the operators in CSR slot order (the same
// row sums in the same order as every SpMV layout, so the same bits).
void NSSolver::apply_matrix_noalloc(int which, const double *x, double *y, int nrhs) {
  if (!scratch_) throw std::runtime_error("apply_matrix_noalloc: no scratch array on this device");
  if (which < 0 || which > 7 || (nrhs != 1 && nrhs != 3) || (which >= 3 && which <= 4 && nrhs != 1))
    throw std::invalid_argument("apply_matrix_noalloc: bad matrix id or nrhs");
  const bool p2 = which <= 2;
  const size_t n = p2 ? (size_t)nu : (size_t)nv;
  const size_t nx = which >= 5 ? (size_t)nv : n * (size_t)nrhs;
  const size_t ny = which >= 5 ? (size_t)nv * 3 : n * (size_t)nrhs;
  const size_t off = (nx + 31) / 32 * 32;  // y starts 256-byte aligned
  if (off + ny > kScratchDoubles)
    throw std::invalid_argument(
        "apply_matrix_noalloc: x and y need " + std::to_string(off + ny) +
        " doubles of scratch, more than the " + std::to_string(kScratchDoubles) +
        " (64 MiB) the allocation-free entry point holds; use cfsa_ns_apply_matrix");
  NoAllocScope no_alloc;
  prepare_operators();
  double *dx = scratch_, *dy = scratch_ + off;
  CFSA_HIP(hipMemcpyAsync(dx, x, nx * sizeof(double), hipMemcpyHostToDevice, stream_));
  if (which >= 5) {
    spmv_grad3(stream_, P_, d_G3.p, dx, dy);
  } else {
    const double *val = which == 0 ? d_M.p : which == 1 ? d_K.p : which == 2 ? d_lhs.p
                        : which == 3 ? d_L.p : d_Lbc.p;
    const DevCsr &A = p2 ? A_ : P_;
    if (nrhs == 3)
      spmv3<double>(stream_, A, val, dx, dy, nullptr);
    else
      spmv1<double>(stream_, A, val, dx, dy);
  }
  CFSA_HIP(hipMemcpyAsync(y, dy, ny * sizeof(double), hipMemcpyDeviceToHost, stream_));
  CFSA_HIP(hipStreamSynchronize(stream_));
}

}  // namespace cfsa
