// lh_noalloc.hip -- max |div u| that allocates nothing: lh_max_div_noalloc.
//
// For callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  The divergence monitor of lh_max_div
// (error, littleharsh.f90:323-352) runs as two kernels on the handle's stream: per-block
// maxima of |div| go into `scratch`, a file-scope device array of kMaxBlocks doubles, and one
// block reduces exactly the entries this call wrote into the caller's device double.  Nothing
// is allocated and nothing needs clearing: the first kernel writes every entry the second
// reads.  The array's device address is looked up once per device, with that device
// current, on the first call there, and kept per device; host code never passes the array's
// name itself.  Every call on a device shares that device's array (it keeps the last call's
// partial maxima), so calls that overlap on one GPU (two streams, two threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/littleharsh_amd.h"
#include "lh_div.hpp"
#include "lh_internal.hpp"

namespace {
constexpr int kBlock = 256;          // columns per block of the first kernel
constexpr int kMaxBlocks = 4096;     // scratch entries: up to 1,048,576 spectral columns
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kMaxBlocks];

double* g_scratch[kMaxDevices] = {};   // device address of `scratch`, per device
std::mutex g_scratch_mu;

// the block maximum of |div| over its columns (NaN kept) -> part[blockIdx.x]
__global__ void __launch_bounds__(kBlock) k_maxdiv_part(LhDivView v, double* part) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  double e = 0.0;
  if (c < v.ncol) e = lh::col_maxdiv(c, v.u1, v.u2, v.u3, v.dthdyu, v.k1x, v.k1z, v.ddthetavi, v.ncol, v.mx, v.nyp);
  __shared__ double red[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) e = lh::nanmax(e, __shfl_down(e, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < kBlock / 64; ++w) m = lh::nanmax(m, red[w]);
    part[blockIdx.x] = m;
  }
}

// one block: the maximum of part[0..n-1] -> *out
__global__ void __launch_bounds__(256) k_maxdiv_final(const double* part, int n, double* out) {
  double e = 0.0;
  for (int q = threadIdx.x; q < n; q += 256) e = lh::nanmax(e, part[q]);
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) e = lh::nanmax(e, __shfl_down(e, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) *out = lh::nanmax(lh::nanmax(red[0], red[1]), lh::nanmax(red[2], red[3]));
}

// the device address of `scratch` on device dev (current), looked up on first use
double* scratch_of(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
    g_scratch[dev] = static_cast<double*>(ptr);
  }
  return g_scratch[dev];
}
}  // namespace

extern "C" int lh_max_div_noalloc(lh_handle* h, double* err_dev) {
  if (!h || !err_dev) return LH_ERR_ARG;
  LhDivView v;
  const int r = lh_div_view(h, &v);
  if (r != LH_OK) return r;
  const int nblk = (v.ncol + kBlock - 1) / kBlock;
  if (nblk > kMaxBlocks)
    return lh_fail(h, LH_ERR_ARG, "lh_max_div_noalloc: more than 1048576 spectral columns (scratch limit)");
  if (v.dev < 0 || v.dev >= kMaxDevices)
    return lh_fail(h, LH_ERR_ARG, "lh_max_div_noalloc: device index beyond 63");
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return lh_fail(h, LH_ERR_HIP, "lh_max_div_noalloc: hipGetDevice failed");
  if (cur != v.dev && hipSetDevice(v.dev) != hipSuccess)
    return lh_fail(h, LH_ERR_HIP, "lh_max_div_noalloc: hipSetDevice failed");
  double* part = scratch_of(v.dev);
  int rc = LH_OK;
  if (!part) {
    rc = lh_fail(h, LH_ERR_HIP, "lh_max_div_noalloc: hipGetSymbolAddress(scratch) failed");
  } else {
    hipLaunchKernelGGL(k_maxdiv_part, dim3(nblk), dim3(kBlock), 0, v.st, v, part);
    hipLaunchKernelGGL(k_maxdiv_final, dim3(1), dim3(256), 0, v.st, (const double*)part, nblk, err_dev);
    if (hipGetLastError() != hipSuccess) rc = lh_fail(h, LH_ERR_HIP, "lh_max_div_noalloc: launch failed");
  }
  if (cur != v.dev) (void)hipSetDevice(cur);
  return rc;
}
