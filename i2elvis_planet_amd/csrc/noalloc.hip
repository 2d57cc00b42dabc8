// noalloc.hip -- i2e_maxvelstep_noalloc: maxvelstep (movemart.c:1225-1262)
// on caller-owned device arrays, for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops).  Its scratch
// (the two velocity maxima) is a file-scope device array instead of a
// context buffer; the call allocates nothing and launches only kernels.
#include <mutex>

#include "ops.hpp"

namespace i2e {

__global__ void k_fill(int n, double v, double *a);  // ctx.hip

// The scratch: |vx|max and |vy|max.  Every GPU the library runs on holds its
// own copy until the process ends; calls on one GPU share it.
constexpr int kScratchDoubles = 4;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// the device address of this device's copy of `scratch`, looked up once per
// device with that device current
int scratch_ptr(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("maxvelstep_noalloc: no current HIP device (or device index >= %d)", kMaxDevices);
    return I2E_ERR_NODEVICE;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("maxvelstep_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
      return I2E_ERR_HIP;
    }
    g_scratch[dev] = (double *)p;
  }
  *out = g_scratch[dev];
  return I2E_OK;
}
}  // namespace

// dt = min(dt, maxxystep dx / max|vx|, maxxystep dy / max|vy|), the same
// operations in the same order as the host side of i2e_maxvelstep
__global__ void k_maxvel_dt(const double *vmax, double xstpx, double ystpy, double maxxystep, double *timestep) {
  if (blockIdx.x || threadIdx.x) return;
  double ts = *timestep, maxvx = vmax[0], maxvy = vmax[1];
  if (maxvx) {
    maxvx = (maxxystep * xstpx) / maxvx;
    ts = maxvx <= ts ? maxvx : ts;
  }
  if (maxvy) {
    maxvy = (maxxystep * ystpy) / maxvy;
    ts = maxvy <= ts ? maxvy : ts;
  }
  *timestep = ts;
}

}  // namespace i2e

extern "C" int i2e_maxvelstep_noalloc(const double *vx_dev, const double *vy_dev, int64_t xnumx, int64_t ynumy,
                                      double xsize, double ysize, double maxxystep, double *timestep_dev,
                                      void *stream) {
  using namespace i2e;
  if (!vx_dev || !vy_dev || !timestep_dev) {
    set_error("maxvelstep_noalloc: null device pointer");
    return I2E_ERR_ARG;
  }
  if (xnumx < 2 || ynumy < 2 || xnumx > 16385 || ynumy > 16385 || xnumx * ynumy > ((int64_t)1 << 28)) {
    set_error("maxvelstep_noalloc: grid %lld x %lld outside [2, 16385]^2 with at most 2^28 nodes",
              (long long)xnumx, (long long)ynumy);
    return I2E_ERR_CAPACITY;
  }
  double *vmax = nullptr;
  ST_OK(scratch_ptr(&vmax));
  hipStream_t s = (hipStream_t)stream;
  // the maxima are combined with atomic max on their bit patterns: start at +0
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, 2, 0.0, vmax);
  hipLaunchKernelGGL(k_maxvel, dim3(512), dim3(256), 0, s, (int)xnumx, (int)ynumy, vx_dev, vy_dev, vmax);
  const double xstpx = xsize / (double)(xnumx - 1), ystpy = ysize / (double)(ynumy - 1);
  hipLaunchKernelGGL(k_maxvel_dt, dim3(1), dim3(64), 0, s, (const double *)vmax, xstpx, ystpy, maxxystep,
                     timestep_dev);
  HIP_OK(hipGetLastError());
  return I2E_OK;
}
