// ops.hip -- FP64 node kernels of the step: the exact (reference-scaled)
// Stokes, gravity and heat operators and right-hand sides, the post-solve
// strain/stress recompute, |g|, maxvelstep and the heat-rate scan.
#include "ops.hpp"

namespace i2e {

// ---------------------------------------------------------------- Stokes
// y = K x for the 3N interleaved system (viterate rows, movemart.c:789-967)
__global__ void k_stokes_apply(SOp<double> o, BcTab bc, const double *__restrict__ x, double *__restrict__ y) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = o.X, Y = o.Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  const VInter v{x, Y};
  auto bcrow = [&](int u) {
    int b = bc.bondm[u];
    double s = x[u];
    for (int n = 0; n < 3; n++) {
      int nb = bc.bondn[3 * b + n];
      if (nb) s -= bc.bondv[4 * b + n + 1] * x[nb - 1];
    }
    return o.mu * s;
  };
  // a velocity BC row, or (coefficient above 1e30, movemart.c:863-879) the
  // continuity row of the cell its first neighbour names
  auto vbc = [&](int u) {
    const int b = bc.bondm[u];
    if (bc.bondv[4 * b + 1] > 1e30) {
      int dm1, dm2;
      cont_shift(bc.bondn[3 * b] - 1 - u, Y, dm1, dm2);
      return o.mu * cont(o, v, i + dm1, j + dm2);
    }
    return bcrow(u);
  };
  double r0, r1, r2;
  if (i && j)
    r0 = bc.bondm[3 * m] ? bcrow(3 * m) : o.mu * cont(o, v, i, j);
  else
    r0 = o.mu * x[3 * m];
  if (j < Y - 1)
    r1 = bc.bondm[3 * m + 1] ? vbc(3 * m + 1) : xmom(o, v, i, j, true);
  else
    r1 = o.mu * x[3 * m + 1];
  if (i < X - 1)
    r2 = bc.bondm[3 * m + 2] ? vbc(3 * m + 2) : ymom(o, v, i, j, true);
  else
    r2 = o.mu * x[3 * m + 2];
  y[3 * m] = r0;
  y[3 * m + 1] = r1;
  y[3 * m + 2] = r2;
}

__global__ void k_stokes_rhs(SOp<double> o, BcTab bc, double *__restrict__ b) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = o.X, Y = o.Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  double r0 = 0, r1 = 0, r2 = 0;
  if (i && j && bc.bondm[3 * m]) r0 = bc.bondv[4 * bc.bondm[3 * m]] * o.mu;
  if (j < Y - 1) {
    int bb = bc.bondm[3 * m + 1];
    if (bb)
      r1 = bc.bondv[4 * bb + 1] > 1e30 ? 0.0 : bc.bondv[4 * bb] * o.mu;  // continuity-type: 0
    else {
      double xkf = (o.gx[i + 1] - o.gx[i - 1]) / 2.0;
      double GX = -(o.gp[m + Y + 1] - o.gp[m + 1]) / xkf;
      r1 = -GX * (o.ro[m] + o.ro[m + 1]) / 2.0;
    }
  }
  if (i < X - 1) {
    int bb = bc.bondm[3 * m + 2];
    if (bb)
      r2 = bc.bondv[4 * bb + 1] > 1e30 ? 0.0 : bc.bondv[4 * bb] * o.mu;
    else {
      double ykf = (o.gy[j + 1] - o.gy[j - 1]) / 2.0;
      double GY = -(o.gp[m + Y + 1] - o.gp[m + Y]) / ykf;
      r2 = -GY * (o.ro[m] + o.ro[m + Y]) / 2.0;
    }
  }
  b[3 * m] = r0;
  b[3 * m + 1] = r1;
  b[3 * m + 2] = r2;
}

// pack (pr, vx, vy) into the interleaved vector and back (movemart.c:979-1012)
__global__ void k_stokes_pack(int X, int Y, const double *pr, const double *vx, const double *vy, double *x) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= Y) return;
  const int m = i * Y + j;
  x[3 * m] = (i && j) ? pr[m] : 0.0;
  x[3 * m + 1] = (j < Y - 1) ? vx[m] : 0.0;
  x[3 * m + 2] = (i < X - 1) ? vy[m] : 0.0;
}

__global__ void k_stokes_unpack(int X, int Y, const double *x, double *pr, double *vx, double *vy) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= Y) return;
  const int m = i * Y + j;
  if (i && j) pr[m] = x[3 * m];
  if (j < Y - 1) vx[m] = x[3 * m + 1];
  if (i < X - 1) vy[m] = x[3 * m + 2];
}

// strain rate / stress recompute, movemart.c:1022-1045 (sxxcalc, syycalc,
// sxycalc with ynval = 0; two-point derivative weights -1/h, 1/h)
__global__ void k_strain(Grid g, NodeF F) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y;
  if (j >= Y || i < 1 || j < 1) return;
  const int m = i * Y + j;
  const double *gx = g.gx, *gy = g.gy;
  double d0, d1, e;
  d1 = 1.0 / (gx[i] - gx[i - 1]);
  d0 = -1.0 / (gx[i] - gx[i - 1]);
  e = 0;
  e += F.vx[(i - 1) * Y + (j - 1)] * d0;
  e += F.vx[i * Y + (j - 1)] * d1;
  F.exx[m] = e;
  F.sxx[m] = e * (2.0 * F.nd[m]);
  d1 = 1.0 / (gy[j] - gy[j - 1]);
  d0 = -1.0 / (gy[j] - gy[j - 1]);
  e = 0;
  e += F.vy[(i - 1) * Y + (j - 1)] * d0;
  e += F.vy[(i - 1) * Y + j] * d1;
  F.eyy[m] = e;
  F.syy[m] = e * (2.0 * F.nd[m]);
  if (i < X - 1 && j < Y - 1) {
    double h = (gy[j] + gy[j + 1]) / 2.0 - (gy[j - 1] + gy[j]) / 2.0;
    e = 0;
    e += F.vx[i * Y + j - 1] * (-1.0 / h) / 2.0;
    e += F.vx[i * Y + j] * (1.0 / h) / 2.0;
    h = (gx[i] + gx[i + 1]) / 2.0 - (gx[i - 1] + gx[i]) / 2.0;
    e += F.vy[(i - 1) * Y + j] * (-1.0 / h) / 2.0;
    e += F.vy[i * Y + j] * (1.0 / h) / 2.0;
    F.exy[m] = e;
    F.sxy[m] = e * (2.0 * F.nu[m]);
  }
}

// maxvelstep reduction (movemart.c:1225-1262): out[0] = max|vx|, out[1] = max|vy|
__global__ void k_maxvel(int X, int Y, const double *vx, const double *vy, double *out) {
  double a = 0, b = 0;
  const int N = X * Y;
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < N; m += gridDim.x * blockDim.x) {
    int i = m / Y, j = m - i * Y;
    if (j < Y - 1) a = fmax(a, fabs(vx[m]));
    if (i < X - 1) b = fmax(b, fabs(vy[m]));
  }
  for (int o = 32; o > 0; o >>= 1) {
    a = fmax(a, __shfl_xor(a, o));
    b = fmax(b, __shfl_xor(b, o));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax((unsigned long long *)&out[0], (unsigned long long)__double_as_longlong(a));
    atomicMax((unsigned long long *)&out[1], (unsigned long long)__double_as_longlong(b));
  }
}

// ---------------------------------------------------------------- gravity
// mean density inside the planet radius (movemart.c:622-651): out += {rosum, ronum}
// per block, reduced on the host in block order for determinism
__global__ void k_grav_mean(Grid g, const double *ro, double R, double *part) {
  const int X = g.X, Y = g.Y, C = (X - 1) * (Y - 1);
  double s = 0, a = 0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
    int i = c / (Y - 1) + 1, j = c % (Y - 1) + 1, m = i * Y + j;
    double dx = (g.gx[i - 1] + g.gx[i] - g.xsize) / 2.0, dy = (g.gy[j - 1] + g.gy[j] - g.ysize) / 2.0;
    double d = pow(dx * dx + dy * dy, 0.5);
    if (d < R) {
      double ar = (g.gx[i] - g.gx[i - 1]) * (g.gy[j] - g.gy[j - 1]);
      s += ar * (ro[m - 1 - Y] + ro[m - Y] + ro[m - 1] + ro[m]) / 4.0;
      a += ar;
    }
  }
  __shared__ double ss[256], sa[256];
  ss[threadIdx.x] = s;
  sa[threadIdx.x] = a;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      ss[threadIdx.x] += ss[threadIdx.x + o];
      sa[threadIdx.x] += sa[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ss[0];
    part[2 * blockIdx.x + 1] = sa[0];
  }
}


// gravity rows (movemart.c:661-749) applied to x (N array over nodes)
__global__ void k_grav_apply(Grid g, GravGeo q, const double *__restrict__ x, double *__restrict__ y) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  double r;
  if (i > 0 && j > 0) {
    if (i == 1 || i == X - 1 || j == 1 || j == Y - 1)
      r = q.mu * x[m];
    else {
      const double *gx = g.gx, *gy = g.gy;
      double dx02 = (gx[i] - gx[i - 2]) / 2.0, dx24 = (gx[i + 1] - gx[i - 1]) / 2.0;
      double dy12 = (gy[j] - gy[j - 2]) / 2.0, dy23 = (gy[j + 1] - gy[j - 1]) / 2.0;
      r = 2.0 / dx02 / (dx02 + dx24) * (x[m - Y] - x[m]) + 2.0 / dx24 / (dx02 + dx24) * (x[m + Y] - x[m]) +
          2.0 / dy12 / (dy12 + dy23) * (x[m - 1] - x[m]) + 2.0 / dy23 / (dy12 + dy23) * (x[m + 1] - x[m]);
      if (grav_dist(g, i, j) > q.R) r += q.mu * 1e+18 * x[m];
    }
  } else
    r = q.mu * x[m];
  y[m] = r;
}

// 1/diagonal of every gravity row (row scaling of the solve)
__global__ void k_grav_rowscale(Grid g, GravGeo q, double *sc) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  double d;
  if (i > 0 && j > 0 && !(i == 1 || i == X - 1 || j == 1 || j == Y - 1)) {
    const double *gx = g.gx, *gy = g.gy;
    double dx02 = (gx[i] - gx[i - 2]) / 2.0, dx24 = (gx[i + 1] - gx[i - 1]) / 2.0;
    double dy12 = (gy[j] - gy[j - 2]) / 2.0, dy23 = (gy[j + 1] - gy[j - 1]) / 2.0;
    d = -2.0 / dx02 / (dx02 + dx24) - 2.0 / dx24 / (dx02 + dx24) - 2.0 / dy12 / (dy12 + dy23) -
        2.0 / dy23 / (dy12 + dy23);
    if (grav_dist(g, i, j) > q.R) d += q.mu * 1e+18;
  } else
    d = q.mu;
  sc[m] = 1.0 / d;
}

__global__ void k_grav_rhs(Grid g, GravGeo q, const double *__restrict__ ro, double *__restrict__ b) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  const double gamma = 6.67384e-11, pival = 2.0 * asin(1.0);
  double r = 0;
  if (i > 0 && j > 0) {
    double d = grav_dist(g, i, j);
    if (i == 1 || i == X - 1 || j == 1 || j == Y - 1)
      r = -q.gxkoef * q.R * q.R / d * q.mu;
    else {
      r = 4.0 / 1.5 * pival * gamma * (ro[m - 1 - Y] + ro[m - Y] + ro[m - 1] + ro[m]) / 4.0;
      if (d > q.R) r -= q.gxkoef * q.R * q.R / d * q.mu * 1e+18;
    }
  }
  b[m] = r;
}

// |g| at basic nodes from gp differences (movemart.c:767-783, keeping the
// reference's y-component expression)
__global__ void k_grav_ga(Grid g, const double *gp, double *ga) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y;
  if (j < 1 || j >= Y - 1 || i < 1 || i >= X - 1) return;
  const int m = i * Y + j;
  double a = (gp[m + Y] + gp[m + Y + 1] - gp[m] - gp[m + 1]) / (g.gx[i + 1] - g.gx[i - 1]);
  double b = (gp[m + 1] + gp[m + Y + 1] - gp[m] - gp[m + 1]) / (g.gy[j + 1] - g.gy[j - 1]);
  ga[m] = pow(a * a + b * b, 0.500);
}

// ---------------------------------------------------------------- heat
__device__ __forceinline__ double qx_v(const Grid &g, const double *tk, const double *kt, int i, int j) {
  const int Y = g.Y;
  double k = (kt[i * Y + j] + kt[(i + 1) * Y + j]) / 2.0, d1 = 1.0 / (g.gx[i + 1] - g.gx[i]), d0 = -1.0 / (g.gx[i + 1] - g.gx[i]);
  double q = 0;
  q += tk[i * Y + j] * d0 * k;
  q += tk[(i + 1) * Y + j] * d1 * k;
  return q;
}
__device__ __forceinline__ double qy_v(const Grid &g, const double *tk, const double *kt, int i, int j) {
  const int Y = g.Y;
  double k = (kt[i * Y + j] + kt[i * Y + j + 1]) / 2.0, d1 = 1.0 / (g.gy[j + 1] - g.gy[j]), d0 = -1.0 / (g.gy[j + 1] - g.gy[j]);
  double q = 0;
  q += tk[i * Y + j] * d0 * k;
  q += tk[i * Y + j + 1] * d1 * k;
  return q;
}

// heaterr(.,0) at every non-BC node (heatmart.c:6-110): sol1 = sources,
// out[0] = max |dT/dt| (interior scan of titerate, heatmart.c:350-377)
__global__ void k_heat_source(const i2e_params *__restrict__ p, Grid g, NodeF F, BcTab bc, double timesum,
                              double *sol1, double *out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y, N = X * Y;
  double rate = 0;
  if (j < Y) {
    const int m = i * Y + j;
    if (!bc.bondm[3 * N + m] && i > 0 && j > 0 && i < X - 1 && j < Y - 1) {
      double ival = F.ht[m] / F.cp[m] / F.ro[m];
      if (p->frictyn > 0 && timesum > p->frictyn && F.ro[m] > 1e+3) {
        double x = g.gx[i], y = g.gy[j];
        int m10 = serch(g.gx, X, x), m20 = serch(g.gy, Y, y);
        W4 w;
        nodewt2(g.gx, g.gy, bn_i0(m10, X), bn_i0(m20, Y), x, y, 0, 0, w);
        double e13 = isum2(w, Y, F.exy, F.sxy);
        nodewt2(g.gx, g.gy, ce_i0(g.gx, m10, x, X), ce_i0(g.gy, m20, y, Y), x, y, -1, -1, w);
        double e14 = isum2(w, Y, F.exx, F.sxx), e15 = isum2(w, Y, F.eyy, F.syy);
        double epsval = 2.0 * e13 + e14 + e15;
        ival += epsval / F.cp[m] / F.ro[m];
      }
      sol1[m] = ival;
      double xkf = (g.gx[i + 1] - g.gx[i - 1]) / 2.0, ykf = (g.gy[j + 1] - g.gy[j - 1]) / 2.0;
      double r = ival + ((qx_v(g, F.tk, F.kt, i, j) - qx_v(g, F.tk, F.kt, i - 1, j)) / xkf +
                         (qy_v(g, F.tk, F.kt, i, j) - qy_v(g, F.tk, F.kt, i, j - 1)) / ykf) /
                            F.cp[m] / F.ro[m];
      rate = fabs(r);
    } else
      sol1[m] = 0;
  }
  for (int o = 32; o > 0; o >>= 1) rate = fmax(rate, __shfl_xor(rate, o));
  if ((threadIdx.x & 63) == 0 && rate > 0)
    atomicMax((unsigned long long *)&out[0], (unsigned long long)__double_as_longlong(rate));
}

// heat rows (heatmart.c:406-432) applied to x
__global__ void k_heat_apply(Grid g, NodeF F, BcTab bc, double dt, const double *__restrict__ x,
                             double *__restrict__ y) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y, N = X * Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  int b = bc.bondm[3 * N + m];
  double r;
  if (!b) {
    double xkf = (g.gx[i + 1] - g.gx[i - 1]) / 2.0, ykf = (g.gy[j + 1] - g.gy[j - 1]) / 2.0;
    r = x[m] / dt - ((qx_v(g, x, F.kt, i, j) - qx_v(g, x, F.kt, i - 1, j)) / xkf +
                     (qy_v(g, x, F.kt, i, j) - qy_v(g, x, F.kt, i, j - 1)) / ykf) /
                        F.cp[m] / F.ro[m];
  } else {
    r = x[m];
    for (int n = 0; n < 3; n++) {
      int nb = bc.bondn[3 * b + n];
      if (nb) r -= bc.bondv[4 * b + n + 1] * x[nb - 1 - 3 * N];
    }
  }
  y[m] = r;
}

// 1/diagonal of every heat row
__global__ void k_heat_rowscale(Grid g, NodeF F, BcTab bc, double dt, double *sc) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y, N = X * Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  if (bc.bondm[3 * N + m]) {
    sc[m] = 1.0;
    return;
  }
  const double *kt = F.kt, *gx = g.gx, *gy = g.gy;
  double xkf = (gx[i + 1] - gx[i - 1]) / 2.0, ykf = (gy[j + 1] - gy[j - 1]) / 2.0;
  double kR = (kt[m] + kt[m + Y]) / 2.0 / (gx[i + 1] - gx[i]), kL = (kt[m - Y] + kt[m]) / 2.0 / (gx[i] - gx[i - 1]);
  double kB = (kt[m] + kt[m + 1]) / 2.0 / (gy[j + 1] - gy[j]), kT = (kt[m - 1] + kt[m]) / 2.0 / (gy[j] - gy[j - 1]);
  sc[m] = 1.0 / (1.0 / dt + ((kR + kL) / xkf + (kB + kT) / ykf) / F.cp[m] / F.ro[m]);
}

__global__ void k_heat_rhs(Grid g, NodeF F, BcTab bc, double dt, const double *sol1, double *rhs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y, N = X * Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  int b = bc.bondm[3 * N + m];
  rhs[m] = b ? bc.bondv[4 * b] : sol1[m] + F.tk[m] / dt;
}

}  // namespace i2e
