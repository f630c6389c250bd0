// The one entry point that allocates nothing: the true relative residual of a system with its scratch (the
// per-block partial sums and the two totals) in a fixed device array of this file instead of the system's own
// buffers. For callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid it); no counterpart in the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "dev_common.hpp"
#include "dfm_internal.hpp"
#include "kernels.hpp"

namespace dfm {
namespace {

// 2 * kMaxPartial partial sums and the two totals: 32,784 bytes per device, held until the process ends.
constexpr int kScratchDoubles = 2 * kMaxPartial + 2;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDev = 64;
std::mutex g_mu;
double* g_ptr[kMaxDev] = {};  // the scratch array's device address, per device

// Address of the scratch array on `dev`, looked up once per device with that device current (the array's name
// is a host-side stand-in, never a device address).
hipError_t scratch_ptr(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDev) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    g_ptr[dev] = (double*)p;
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

}  // namespace
}  // namespace dfm

using namespace dfm;

extern "C" int dfm_sys_true_residual_noalloc(dfm_sys sys, double* rel) {
  DFM_CHECK(sys && rel, DFM_ERR_ARG, "dfm_sys_true_residual_noalloc: null argument");
  Sys& s = sys->s;
  DFM_CHECK(s.n > 0 && (long long)s.n * s.len <= 0x7FFFFFFFLL, DFM_ERR_ARG,
            "dfm_sys_true_residual_noalloc: the system has more dofs than the kernel indexes (2^31 - 1)");
  DFM_CHECK(s.h_pinned_d != nullptr, DFM_ERR_STATE, "dfm_sys_true_residual_noalloc: system without host mirror");
  double* sc = nullptr;
  DFM_HIP(scratch_ptr(s.ctx->device, &sc));
  hipStream_t st = s.ctx->stream;
  // every partial slot of the launch is written before it is read: nothing to clear
  DFM_HIP(launch_true_residual(s.n, s.len, spmv_tpr(s.ncrs, s.n), s.d_rowptr.p, s.d_colind.p, s.d_crs.p, s.d_dia.p,
                               s.d_upd.p, s.d_rhs.p, sc, sc + 2 * kMaxPartial, st));
  DFM_HIP(hipMemcpyAsync(s.h_pinned_d, sc + 2 * kMaxPartial, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  DFM_HIP(hipStreamSynchronize(st));
  *rel = s.h_pinned_d[1] > 0 ? std::sqrt(s.h_pinned_d[0] / s.h_pinned_d[1]) : 0.0;
  return DFM_OK;
}
