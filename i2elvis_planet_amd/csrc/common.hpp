// common.hpp -- shared device helpers of the I2ELVIS step library (gfx950).
//
// Grid geometry, the reference's cell search and two-point Fornberg weights
// (markmart.c:1051-1088, movemart.c:1269-1319) with the operation order that
// makes marker arithmetic bit-identical to the serial C code when both sides
// are compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "../../include/i2e.h"

#define I2E_HD __host__ __device__ __forceinline__

namespace i2e {

void set_error(const char *fmt, ...);

#define HIP_OK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ::i2e::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return I2E_ERR_HIP;                                                             \
    }                                                                                 \
  } while (0)

#define ST_OK(call)              \
  do {                           \
    int s_ = (call);             \
    if (s_ != I2E_OK) return s_; \
  } while (0)

// Per-launch HIP-event timing of one kernel class on its launch stream, for
// roofline figures measured inside a normal run.  Host-side only: launch
// sites bracket the launch (or a group of launches forming one unit of work)
// with begin/end and state two byte counts: the algorithmic bytes of the
// unit (each distinct element of the data it must touch counted once: 4 B
// per float32 marker field, 1 B per type, 8 B per FP64 node value) and the
// representation bytes of this implementation (e.g. the explicit Galerkin
// stencils a smoothing pass reads in place of matrix-free rows).
enum KId {
  K_ADVECT = 0,
  K_P2G,
  K_GSDOT,
  K_GSAXPY,
  K_VSMOOTH0,
  K_VSMOOTHC,
  K_STOKES_APPLY,
  K_PRECOND,
  K_HEATMARK,
  K_VCOARSE,
  K_COUNT
};

// Every `every`-th launch of the target is bracketed (a sample: the event
// pairs themselves add latency to the stream they are recorded on).
struct KTimer {
  int target = -1, every = 1;
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  int64_t launches = 0, seen = 0;
  double bytes = 0, repr = 0;
  bool open = false;
  void begin(int id, hipStream_t s) {
    if (id != target || used + 2 > ev.size()) return;
    open = (seen++ % every) == 0;
    if (open) (void)hipEventRecord(ev[used], s);
  }
  void end(int id, hipStream_t s, double alg, double rep) {
    if (id != target || !open) return;
    open = false;
    (void)hipEventRecord(ev[used + 1], s);
    used += 2;
    launches++;
    bytes += alg;
    repr += rep;
  }
  // a single-kernel unit: the events come back to be given to the launch
  // itself (hipExtLaunchKernelGGL), which stamps them from the dispatch's
  // own begin and end -- the kernel's duration as a kernel trace reports
  // it, without the ~3 us an event pair around the launch adds
  bool take(int id, double alg, double rep, hipEvent_t *e0, hipEvent_t *e1) {
    if (id != target || used + 2 > ev.size()) return false;
    if ((seen++ % every) != 0) return false;
    *e0 = ev[used];
    *e1 = ev[used + 1];
    used += 2;
    launches++;
    bytes += alg;
    repr += rep;
    return true;
  }
};

// launch KERNEL on STREAM, timed by the KTimer* KT as unit ID when it samples
// this launch (ALG / REP: its algorithmic and representation bytes)
#define KT_LAUNCH(KT, ID, ALG, REP, KERNEL, GRID, BLOCK, SHM, STREAM, ...)                              \
  do {                                                                                                  \
    hipEvent_t kt_e0_, kt_e1_;                                                                          \
    if ((KT) && (KT)->take((ID), (ALG), (REP), &kt_e0_, &kt_e1_))                                       \
      hipExtLaunchKernelGGL(KERNEL, GRID, BLOCK, SHM, STREAM, kt_e0_, kt_e1_, 0, __VA_ARGS__);          \
    else                                                                                                \
      hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHM, STREAM, __VA_ARGS__);                                \
  } while (0)

// v of the lane that the DPP control selects, both 32-bit halves moved by the
// VALU's data-parallel primitives; 0 in rows outside ROW_MASK and where the
// source lane lies outside the lane's row (row shifts)
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_get(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add(double v) {
  return v + dpp_get<CTRL, ROW_MASK>(v);
}

// Sum over the 64 lanes of a wavefront, returned to every lane; all lanes
// must be active.  Fixed order (reproducible): pairs and quads by quad
// permutes, 8- and 16-lane rows by half-row and row mirrors, then rows
// chained by row broadcasts 15 and 31, so lane 63 ends with the total.  The
// whole reduction stays in the VALU (lane shuffles through ds_bpermute cost
// an LDS round trip per step, which bounded the Gram-Schmidt projection).
__device__ __forceinline__ double wave_sum(double v) {
  v = dpp_add<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x141, 0xF>(v);  // row_half_mirror
  v = dpp_add<0x140, 0xF>(v);  // row_mirror
  v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_add<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ void aadd(double *a, double v) { unsafeAtomicAdd(a, v); }

// Segmented sums over the lanes of a wavefront whose keys are equal on
// contiguous runs (the scatters' cell index): sum(v) is the inclusive prefix
// of v within the lane's run, so the run's last lane (tail) holds its total.
// Hillis-Steele in DPP: row shifts 1, 2, 4, 8, then row broadcasts 15 and 31,
// each step adding the shifted value only where its source lane lies in the
// same run (lane - k >= the run's first lane, fixed once per wavefront for
// every sum).  The step weights are 0/1 factors of an FMA, not a select: a
// select lets the compiler run the DPP move under the predicate's exec mask,
// and DPP reads lanes that EXEC disables as out of range (0).  All lanes must
// be active; the summed values must be finite.
struct SegScan {
  double f1, f2, f4, f8, f15, f31;
  bool tail;
  __device__ __forceinline__ SegScan(int key, int lane) {
    const int kprev = __shfl_up(key, 1, 64), knext = __shfl_down(key, 1, 64);
    const bool head = lane == 0 || kprev != key;
    tail = lane == 63 || knext != key;
    const unsigned long long heads = __ballot(head);
    const int s = 63 - __clzll(heads & ((2ull << lane) - 1ull));  // first lane of the run
    f1 = lane - 1 >= s ? 1.0 : 0.0;
    f2 = lane - 2 >= s ? 1.0 : 0.0;
    f4 = lane - 4 >= s ? 1.0 : 0.0;
    f8 = lane - 8 >= s ? 1.0 : 0.0;
    f15 = 16 * (lane >> 4) - 1 >= s ? 1.0 : 0.0;
    f31 = 31 >= s ? 1.0 : 0.0;
  }
  __device__ __forceinline__ double sum(double v) const {
    v = __builtin_fma(dpp_get<0x111, 0xF>(v), f1, v);   // row_shr:1
    v = __builtin_fma(dpp_get<0x112, 0xF>(v), f2, v);   // row_shr:2
    v = __builtin_fma(dpp_get<0x114, 0xF>(v), f4, v);   // row_shr:4
    v = __builtin_fma(dpp_get<0x118, 0xF>(v), f8, v);   // row_shr:8
    v = __builtin_fma(dpp_get<0x142, 0xA>(v), f15, v);  // row_bcast:15 into rows 1 and 3
    v = __builtin_fma(dpp_get<0x143, 0xC>(v), f31, v);  // row_bcast:31 into rows 2 and 3
    return v;
  }
};

// Grid lines of one level.  Node m = m1*Y + m2.
struct Grid {
  int X, Y;
  double xsize, ysize;
  const double *gx, *gy;  // device
};

// m1serch / m2serch (markmart.c:1051-1088): bisection over grid lines,
// clamped to n-2, i.e. the largest m <= n-2 with g[m] <= x (0 below g[0]).
// Inside the grid the same m is reached from the cell that the mean spacing
// predicts, stepping down while g[m] > x and up while g[m+1] <= x (grid
// lines strictly increasing): one or two loads on a near-uniform grid
// instead of a chain of log2(n) dependent ones.  Positions outside [g[0],
// g[n-1]) and NaN take the bisection itself.
I2E_HD int serch(const double *g, int n, double x) {
  const double g0 = g[0], gn = g[n - 1];
  const double f = (x - g0) / (gn - g0) * (double)(n - 1);
  if (x >= g0 && x < gn && f >= 0 && f < (double)(n - 1)) {
    int m = (int)f;
    if (m > n - 2) m = n - 2;
    while (m > 0 && g[m] > x) m--;
    while (m < n - 2 && g[m + 1] <= x) m++;
    return m;
  }
  int m, m0 = 0, m1 = n - 1;
  do {
    m = (m0 + m1) / 2;
    if (g[m] > x)
      m1 = m;
    else
      m0 = m;
  } while ((m1 - m0) > 1);
  if (m0 > n - 2) m0 = n - 2;
  return m0;
}

// fdweight(n=1, m=0): interpolation weights at xi between x0 and x1.
I2E_HD void fd2_interp(double x0, double x1, double xi, double &w0, double &w1) {
  double c5 = x0 - xi, c4 = x1 - xi, c3 = x1 - x0;
  w1 = -1.0 * c5 * 1.0 / c3;
  w0 = c4 * 1.0 / c3;
}

// fdweight(n=1, m=1): two-point derivative weights -1/h, 1/h.
I2E_HD void fd2_deriv(double x0, double x1, double &d0, double &d1) {
  const double c3 = x1 - x0;
  d1 = 1.0 / c3;
  d0 = -1.0 / c3;
}

I2E_HD double subline(const double *g, int m, int sh) {
  return sh < 0 ? (g[m] + g[m - 1]) / 2.0 : (sh > 0 ? (g[m] + g[m + 1]) / 2.0 : g[m]);
}

// Bilinear stencil (origin i0,j0 + weights) on one staggered sub-grid.
struct W4 {
  int i0, j0;
  double wx0, wx1, wy0, wy1;
};

I2E_HD void nodewt2(const double *gx, const double *gy, int i0, int j0, double x, double y, int ynx, int yny, W4 &w) {
  w.i0 = i0;
  w.j0 = j0;
  fd2_interp(subline(gx, i0, ynx), subline(gx, i0 + 1, ynx), x, w.wx0, w.wx1);
  fd2_interp(subline(gy, j0, yny), subline(gy, j0 + 1, yny), y, w.wy0, w.wy1);
}

// m1 outer, m2 inner, from zero (markmart.c:1415-1422)
I2E_HD double isum(const W4 &w, int Y, const double *f) {
  int m = w.i0 * Y + w.j0;
  double e = 0;
  e += (w.wx0 * w.wy0) * f[m];
  e += (w.wx0 * w.wy1) * f[m + 1];
  e += (w.wx1 * w.wy0) * f[m + Y];
  e += (w.wx1 * w.wy1) * f[m + Y + 1];
  return e;
}

I2E_HD double isum2(const W4 &w, int Y, const double *f, const double *h) {
  int m = w.i0 * Y + w.j0;
  double e = 0;
  e += (w.wx0 * w.wy0) * f[m] * h[m];
  e += (w.wx0 * w.wy1) * f[m + 1] * h[m + 1];
  e += (w.wx1 * w.wy0) * f[m + Y] * h[m + Y];
  e += (w.wx1 * w.wy1) * f[m + Y + 1] * h[m + Y + 1];
  return e;
}

// stencil origins of the staggered sub-grids (allinterv/allintere,
// markmart.c:1378-1543; intermod = 0)
I2E_HD int vx_j0(const double *gy, int Y, int m20, double y) {
  int j = m20;
  if (y < (gy[m20] + gy[m20 + 1]) / 2.0) j -= 1;
  if (j < 0) j = 0;
  if (j > Y - 3) j = Y - 3;
  return j;
}
I2E_HD int vy_i0(const double *gx, int X, int m10, double x) {
  int i = m10;
  if (x < (gx[m10] + gx[m10 + 1]) / 2.0) i -= 1;
  if (i < 0) i = 0;
  if (i > X - 3) i = X - 3;
  return i;
}
I2E_HD int bn_i0(int m10, int n) {
  int i = m10;
  if (i < 1) i = 1;
  if (i > n - 3) i = n - 3;
  return i;
}
I2E_HD int ce_i0(const double *g, int m10, double x, int n) {
  int i = m10;
  if (x > (g[m10] + g[m10 + 1]) / 2.0) i += 1;
  if (i < 1) i = 1;
  if (i > n - 2) i = n - 2;
  return i;
}

I2E_HD double dabs(double a) { return a >= 0 ? a : -a; }

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace i2e
