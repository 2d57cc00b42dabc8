// rows.hpp -- matrix-free rows of the staggered Stokes operator, templated on
// the arithmetic type and on how unknowns are read.
//
// The rows are those viterate assembles for gausmat4 (movemart.c:4-279 x/y
// Stokes with the Duretz et al. (2011) free-surface stabilisation,
// movemart.c:287-426 continuity, sxxcalc/syycalc/sxycalc 1326-1649 with
// stoksfd = 0).  The same code evaluates the FP64 residual on the finest grid
// and the FP32/FP64 multigrid operators on the coarse grids (there with
// restricted viscosity, density and potential fields).
#pragma once
#include "common.hpp"

namespace i2e {

template <class T>
struct SOp {
  int X, Y;
  const double *gx, *gy;       // grid lines of this level
  const T *nu, *nd, *ro, *gp;  // basic-node and cell fields of this level
  double dt, mu;               // timestep (stabilisation), row scale mukoef
};

// unknown readers ---------------------------------------------------------
template <class T>
struct VSplit {  // separate component arrays, index m1*Y+m2
  const T *x, *y, *p;
  int Y;
  __device__ __forceinline__ T vx(int i, int j) const { return x[i * Y + j]; }
  __device__ __forceinline__ T vy(int i, int j) const { return y[i * Y + j]; }
  __device__ __forceinline__ T pp(int i, int j) const { return p ? p[i * Y + j] : T(0); }
};
struct VInter {  // the reference's interleaved vector, 3m + {0:P,1:Vx,2:Vy}
  const double *x;
  int Y;
  __device__ __forceinline__ double vx(int i, int j) const { return x[3 * (i * Y + j) + 1]; }
  __device__ __forceinline__ double vy(int i, int j) const { return x[3 * (i * Y + j) + 2]; }
  __device__ __forceinline__ double pp(int i, int j) const { return x[3 * (i * Y + j) + 0]; }
};
template <class T>
struct VUnit {  // unit vector e_(comp, idx): extracts one matrix column
  int comp, idx, Y;
  __device__ __forceinline__ T vx(int i, int j) const { return (comp == 1 && i * Y + j == idx) ? T(1) : T(0); }
  __device__ __forceinline__ T vy(int i, int j) const { return (comp == 2 && i * Y + j == idx) ? T(1) : T(0); }
  __device__ __forceinline__ T pp(int i, int j) const { return (comp == 0 && i * Y + j == idx) ? T(1) : T(0); }
};

// The row functions take the operator O by type: SOp<T> for the solvers, or
// a caller's struct with the same members whose field members return a
// dual number (the adjoint's field derivatives).

// sigma_xy at basic node (i,j) (sxycalc, movemart.c:1515-1649, stencil
// clamped at the edges exactly as the reference clamps it)
template <class T, class O, class A>
__device__ __forceinline__ T sxy_at_g(const O &o, const A &v, int i, int j) {
  T a = 0, b = 0;
  int jmin = j - 1 < 0 ? 0 : j - 1, jmax = j > o.Y - 2 ? o.Y - 2 : j;
  if (jmax > jmin) {
    double h = (o.gy[jmax] + o.gy[jmax + 1]) / 2.0 - (o.gy[jmin] + o.gy[jmin + 1]) / 2.0;
    a = (v.vx(i, jmax) - v.vx(i, jmin)) / T(h);
  }
  int imin = i - 1 < 0 ? 0 : i - 1, imax = i > o.X - 2 ? o.X - 2 : i;
  if (imax > imin) {
    double h = (o.gx[imax] + o.gx[imax + 1]) / 2.0 - (o.gx[imin] + o.gx[imin + 1]) / 2.0;
    b = (v.vy(imax, j) - v.vy(imin, j)) / T(h);
  }
  return o.nu[i * o.Y + j] * (a + b);
}

// x-Stokes row at Vx(i,j), i in [1,X-2], j in [0,Y-2] (xstokserr)
template <class T, class O, class A>
__device__ __forceinline__ T xmom_g(const O &o, const A &v, int i, int j, bool withP) {
  const int Y = o.Y;
  const double *gx = o.gx, *gy = o.gy;
  const T xkf = T((gx[i + 1] - gx[i - 1]) / 2.0), ykf = T(gy[j + 1] - gy[j]);
  const int v0 = i * Y + j, v1 = v0 + 1, v2 = v0 + Y, v3 = v2 + 1;
  T sL = T(2) * o.nd[v1] * (v.vx(i, j) - v.vx(i - 1, j)) / T(gx[i] - gx[i - 1]);
  T sR = T(2) * o.nd[v3] * (v.vx(i + 1, j) - v.vx(i, j)) / T(gx[i + 1] - gx[i]);
  T sT = sxy_at_g<T>(o, v, i, j), sB = sxy_at_g<T>(o, v, i, j + 1);
  T r = (sR - sL) / xkf + (sB - sT) / ykf;
  if (withP) r -= (v.pp(i + 1, j + 1) - v.pp(i, j + 1)) / xkf;
  // -gx dt/2 (vx drho/dx + vy drho/dy), movemart.c:116-125
  T GX = -(o.gp[v3] - o.gp[v1]) / xkf;
  T c = T(-o.dt / 2.0) * GX;
  T dxx = T(gx[i + 1] - gx[i - 1]);
  T dro = o.ro[v1] - o.ro[v0];
  T ca = c * T(gx[i] - gx[i - 1]) / dxx / T(2) * dro / ykf;
  T cb = c * T(gx[i + 1] - gx[i]) / dxx / T(2) * dro / ykf;
  r += c * (o.ro[v2] + o.ro[v3] - o.ro[v0 - Y] - o.ro[v1 - Y]) / dxx / T(2) * v.vx(i, j);
  r += ca * (v.vy(i, j) + v.vy(i, j + 1)) + cb * (v.vy(i - 1, j) + v.vy(i - 1, j + 1));
  return r;
}

// y-Stokes row at Vy(i,j), i in [0,X-2], j in [1,Y-2] (ystokserr)
template <class T, class O, class A>
__device__ __forceinline__ T ymom_g(const O &o, const A &v, int i, int j, bool withP) {
  const int Y = o.Y;
  const double *gx = o.gx, *gy = o.gy;
  const T xkf = T(gx[i + 1] - gx[i]), ykf = T((gy[j + 1] - gy[j - 1]) / 2.0);
  const int v0 = i * Y + j, v1 = v0 + 1, v2 = v0 + Y, v3 = v2 + 1;
  T sT = T(2) * o.nd[v2] * (v.vy(i, j) - v.vy(i, j - 1)) / T(gy[j] - gy[j - 1]);
  T sB = T(2) * o.nd[v3] * (v.vy(i, j + 1) - v.vy(i, j)) / T(gy[j + 1] - gy[j]);
  T sLx = sxy_at_g<T>(o, v, i, j), sRx = sxy_at_g<T>(o, v, i + 1, j);
  T r = (sB - sT) / ykf + (sRx - sLx) / xkf;
  if (withP) r -= (v.pp(i + 1, j + 1) - v.pp(i + 1, j)) / ykf;
  T GY = -(o.gp[v3] - o.gp[v2]) / ykf;
  T c = T(-o.dt / 2.0) * GY;
  T dyy = T(gy[j + 1] - gy[j - 1]);
  T dro = o.ro[v2] - o.ro[v0];
  T ca = c * T(gy[j] - gy[j - 1]) / dyy / T(2) * dro / xkf;
  T cb = c * T(gy[j + 1] - gy[j]) / dyy / T(2) * dro / xkf;
  r += c * (o.ro[v1] + o.ro[v3] - o.ro[v0 - 1] - o.ro[v2 - 1]) / dyy / T(2) * v.vy(i, j);
  r += ca * (v.vx(i, j) + v.vx(i + 1, j)) + cb * (v.vx(i, j - 1) + v.vx(i + 1, j - 1));
  return r;
}

// continuity at cell (i,j), i,j >= 1, unscaled (conterr)
template <class T, class O, class A>
__device__ __forceinline__ T cont_g(const O &o, const A &v, int i, int j) {
  return (v.vx(i, j - 1) - v.vx(i - 1, j - 1)) / T(o.gx[i] - o.gx[i - 1]) +
         (v.vy(i - 1, j) - v.vy(i - 1, j - 1)) / T(o.gy[j] - o.gy[j - 1]);
}

template <class T, class A>
__device__ __forceinline__ T sxy_at(const SOp<T> &o, const A &v, int i, int j) {
  return sxy_at_g<T>(o, v, i, j);
}
template <class T, class A>
__device__ __forceinline__ T xmom(const SOp<T> &o, const A &v, int i, int j, bool withP) {
  return xmom_g<T>(o, v, i, j, withP);
}
template <class T, class A>
__device__ __forceinline__ T ymom(const SOp<T> &o, const A &v, int i, int j, bool withP) {
  return ymom_g<T>(o, v, i, j, withP);
}
template <class T, class A>
__device__ __forceinline__ T cont(const SOp<T> &o, const A &v, int i, int j) {
  return cont_g<T>(o, v, i, j);
}

// ---------------------------------------------------------------------------
// scalar vertex-grid operator used by the heat and gravity preconditioners:
//   c u - [ (kR (u+ - u)/hR - kL (u - u-)/hL) / ((hL+hR)/2) + (same in y) ]
// with k averaged arithmetically between neighbours (qxcalc/qycalc,
// heatmart.c:162-312; the gravity rows movemart.c:700-728 are k = 1, c = 0).
template <class T>
struct POp {
  int X, Y;
  const double *x, *y;  // point coordinates
  const T *k, *c;
  const int8_t *mask;   // 1: row solved pointwise (u = r / dm)
  const T *dm;
};

template <class T, class F>
__device__ __forceinline__ T prow(const POp<T> &o, const F &u, int i, int j) {
  const int Y = o.Y, m = i * Y + j;
  if (o.mask[m]) return o.dm[m] * u(i, j);
  const T uc = u(i, j), kc = o.k[m];
  T r = o.c[m] * uc;
  if (i > 0 && i < o.X - 1) {
    double hL = o.x[i] - o.x[i - 1], hR = o.x[i + 1] - o.x[i];
    T kL = (o.k[m - Y] + kc) / T(2), kR = (o.k[m + Y] + kc) / T(2);
    r -= (kR * (u(i + 1, j) - uc) / T(hR) - kL * (uc - u(i - 1, j)) / T(hL)) / T((hL + hR) / 2.0);
  }
  if (j > 0 && j < Y - 1) {
    double hT = o.y[j] - o.y[j - 1], hB = o.y[j + 1] - o.y[j];
    T kT = (o.k[m - 1] + kc) / T(2), kB = (o.k[m + 1] + kc) / T(2);
    r -= (kB * (u(i, j + 1) - uc) / T(hB) - kT * (uc - u(i, j - 1)) / T(hT)) / T((hT + hB) / 2.0);
  }
  return r;
}

}  // namespace i2e
