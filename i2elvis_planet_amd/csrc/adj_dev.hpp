// adj_dev.hpp -- device helpers shared by the flow adjoint and the coupled
// peak-temperature adjoint: the RK4 velocity of movemark with its
// interpolation stencils kept for the reverse, and the tent weights of the
// viscosity scatter (markmart.c:74-97, 840-911, 1378-1453).
#pragma once
#include "ctx.hpp"
#include "kernels.hpp"
#include "ops.hpp"

namespace i2e {

// ------------------------------------------------------------------ RK4 and its reverse
// allinterv at (x, y) with the stencils and weights kept for the reverse
struct Interp {
  W4 wx, wy;           // vx and vy sub-grid stencils
  double cx, cy;       // 1 where clampxy left the coordinate as it was
  double ivx, ivy;     // interpolated velocity
};
__device__ __forceinline__ void interp_at(const Grid &g, const double *vx, const double *vy, double x, double y,
                                          Interp &I) {
  I.cx = (x >= 0 && x <= g.xsize) ? 1.0 : 0.0;
  I.cy = (y >= 0 && y <= g.ysize) ? 1.0 : 0.0;
  if (x < 0)
    x = 0;
  else if (x > g.xsize)
    x = g.xsize;
  if (y < 0)
    y = 0;
  else if (y > g.ysize)
    y = g.ysize;
  const int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
  nodewt2(g.gx, g.gy, m10, vx_j0(g.gy, g.Y, m20, y), x, y, 0, +1, I.wx);
  I.ivx = isum(I.wx, g.Y, vx);
  nodewt2(g.gx, g.gy, vy_i0(g.gx, g.X, m10, x), m20, x, y, +1, 0, I.wy);
  I.ivy = isum(I.wy, g.Y, vy);
}
// d(ivx, ivy)/d(x, y) transposed onto (bx, by): the bilinear weights are
// linear in each coordinate inside the stencil's cell
__device__ __forceinline__ void interp_grad_t(const Grid &g, const W4 &w, int ynx, int yny, const double *f,
                                              double bar, double cx, double cy, double &px, double &py) {
  const int Y = g.Y, m = w.i0 * Y + w.j0;
  const double hx = subline(g.gx, w.i0 + 1, ynx) - subline(g.gx, w.i0, ynx);
  const double hy = subline(g.gy, w.j0 + 1, yny) - subline(g.gy, w.j0, yny);
  const double f00 = f[m], f01 = f[m + 1], f10 = f[m + Y], f11 = f[m + Y + 1];
  const double dfx = (w.wy0 * (f10 - f00) + w.wy1 * (f11 - f01)) / hx;
  const double dfy = (w.wx0 * (f01 - f00) + w.wx1 * (f11 - f10)) / hy;
  px += bar * dfx * cx;
  py += bar * dfy * cy;
}
__device__ __forceinline__ void scatter_w(const W4 &w, int Y, double bar, double *dst, int comp) {
  if (bar == 0) return;
  const int m = w.i0 * Y + w.j0;
  aadd(&dst[3 * m + comp], w.wx0 * w.wy0 * bar);
  aadd(&dst[3 * (m + 1) + comp], w.wx0 * w.wy1 * bar);
  aadd(&dst[3 * (m + Y) + comp], w.wx1 * w.wy0 * bar);
  aadd(&dst[3 * (m + Y + 1) + comp], w.wx1 * w.wy1 * bar);
}

struct Rk4 {
  Interp s[4];
  double vx, vy;
};
// movemark's velocity (k_movemark): RK4 stages, or one evaluation for Euler
__device__ __forceinline__ void rk4(const i2e_params *p, const Grid &g, const double *vx, const double *vy,
                                    double x0, double y0, double dt, Rk4 &R) {
  interp_at(g, vx, vy, x0, y0, R.s[0]);
  if (p->markmod == 1) {
    R.vx = R.s[0].ivx;
    R.vy = R.s[0].ivy;
    return;
  }
  interp_at(g, vx, vy, x0 + R.s[0].ivx * dt / 2.0, y0 + R.s[0].ivy * dt / 2.0, R.s[1]);
  interp_at(g, vx, vy, x0 + R.s[1].ivx * dt / 2.0, y0 + R.s[1].ivy * dt / 2.0, R.s[2]);
  interp_at(g, vx, vy, x0 + R.s[2].ivx * dt, y0 + R.s[2].ivy * dt, R.s[3]);
  R.vx = (R.s[0].ivx + 2.0 * R.s[1].ivx + 2.0 * R.s[2].ivx + R.s[3].ivx) / 6.0;
  R.vy = (R.s[0].ivy + 2.0 * R.s[1].ivy + 2.0 * R.s[2].ivy + R.s[3].ivy) / 6.0;
}

// the two tent weights of the marker (nearest basic node, cell centre)
struct Tent {
  int nn, nc;      // node indices, -1 none
  double wn, wc;
};
__device__ __forceinline__ Tent tent(const Grid &g, double x, double y) {
  const int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
  W4 w;
  nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
  Tent T{-1, -1, 0.0, 0.0};
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      const double ddx = a ? w.wx1 : w.wx0, ddy = b ? w.wy1 : w.wy0;
      if (ddx > 0.5 && ddy > 0.5) {
        T.nn = (m10 + a) * g.Y + m20 + b;
        T.wn = 4.0 * (ddx - 0.5) * (ddy - 0.5);
      }
    }
  const double ddx = 1.0 - dabs(w.wx0 - 0.5), ddy = 1.0 - dabs(w.wy0 - 0.5);
  if (ddx > 0.5 && ddy > 0.5) {
    T.nc = (m10 + 1) * g.Y + m20 + 1;
    T.wc = 4.0 * (ddx - 0.5) * (ddy - 0.5);
  }
  return T;
}

// S = {sum w eta, sum w} for nu (basic nodes) and nd (cell centres)

// flow_adjoint.hip: dJ/d(nu, nd, rho, gp) of the Stokes rows for the
// multipliers lam of the solution x, and the transposed solve K^T mu = g by
// FGMRES with the forward preconditioner of the context's current set-up (g
// is overwritten; scratch holds three 3N vectors)
__global__ void k_stokes_field_adj(SOp<double> o, BcTab bc, const double *__restrict__ x,
                                   const double *__restrict__ lam, double *lnu, double *lnd, double *lro,
                                   double *lgp);
int stokes_solve_t(i2e_ctx *c, double *g, double *mu, double *scratch, int *its, double *rr);

}  // namespace i2e
