// flow_adjoint.hip -- discrete adjoint of one Stokes solve and the marker
// advection that reads it: the transposed Stokes operator, the adjoint solve
// on it with the forward solver's preconditioner, and the reversed RK4
// interpolation.
//
// Objective (a sensitivity of the iron sinking that core formation starts
// from; no counterpart in the reference):
//   J = sum_i m_i (v_i . r_i) / sum_i m_i
// over the iron markers (is_iron) inside the grid with markk > 0, m_i the
// type's density markro, r_i the unit vector from the box centre and v_i the
// velocity movemark would move marker i with in this step (RK4, or Euler for
// markmod 1; markmart.c:74-97) from the last Stokes solution, with the
// context's current time step.  Parameter: s, a common factor on every
// rock's activation energy (markdh), at s = 1.
//
// Chain differentiated, newest first:
//   RK4 stages (allinterv, markmart.c:1378-1453): dJ/d(node vx, vy) by the
//     transposed interpolation, the stage positions moving with the stage
//     velocities (their bilinear weights differentiated inside a cell);
//   the Stokes solve K x = b (movemart.c:4-426, gausmat4): K^T mu = dJ/dx by
//     FGMRES on P^-1 K^T P with the forward preconditioner, P = diag(1 on
//     velocity rows, 1/mukoef on pressure rows): the continuity rows are
//     mukoef times the transposed gradient rows, so P^-1 K^T P = K where the
//     viscous block is symmetric (stabilisation and BC rows make it nearly
//     so), and the residual keeps the forward rows' scaling;
//     then dJ/d(node field) = -mu^T d(K x - b)/d(field) for the shear and
//     normal viscosities (nu, nd), the density and the gravity potential,
//     evaluated by the same row code in dual numbers;
//   the viscosity scatter of ronurecalc (markmart.c:840-911, node
//     normalisation and the [nubeg, nuend] clamp :917-1009): the tent
//     weights of the marker's nearest basic node and cell centre;
//   viscalc's Arrhenius branch (markmart.c:1855-2063): d eta / d s =
//     eta dh / (R T) where no exponent cap or viscosity limit holds.
// Held fixed: the node fields other than the viscosities (density and
// gravity do not depend on s), marker positions and temperatures, and the
// time steps (the stabilisation's from the solve, the RK4's from the
// caller).
#include <string.h>

#include "adj_dev.hpp"
#include "viscalc.hpp"

namespace i2e {

// ------------------------------------------------------------------ dual numbers
// value + four partial derivatives (nu, nd, rho, gp of one node)
struct D4 {
  double v, d[4];
  __device__ __forceinline__ D4() {}
  __device__ __forceinline__ D4(double a) : v(a) { d[0] = d[1] = d[2] = d[3] = 0.0; }
};
__device__ __forceinline__ D4 operator+(const D4 &a, const D4 &b) {
  D4 r(a.v + b.v);
#pragma unroll
  for (int k = 0; k < 4; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ __forceinline__ D4 operator-(const D4 &a, const D4 &b) {
  D4 r(a.v - b.v);
#pragma unroll
  for (int k = 0; k < 4; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
__device__ __forceinline__ D4 operator-(const D4 &a) {
  D4 r(-a.v);
#pragma unroll
  for (int k = 0; k < 4; k++) r.d[k] = -a.d[k];
  return r;
}
__device__ __forceinline__ D4 operator*(const D4 &a, const D4 &b) {
  D4 r(a.v * b.v);
#pragma unroll
  for (int k = 0; k < 4; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ D4 operator/(const D4 &a, const D4 &b) {
  const double q = a.v / b.v;
  D4 r(q);
#pragma unroll
  for (int k = 0; k < 4; k++) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
  return r;
}
__device__ __forceinline__ D4 operator+(const D4 &a, double b) { return a + D4(b); }
__device__ __forceinline__ D4 operator+(double a, const D4 &b) { return D4(a) + b; }
__device__ __forceinline__ D4 operator-(const D4 &a, double b) { return a - D4(b); }
__device__ __forceinline__ D4 operator-(double a, const D4 &b) { return D4(a) - b; }
__device__ __forceinline__ D4 operator*(const D4 &a, double b) { return a * D4(b); }
__device__ __forceinline__ D4 operator*(double a, const D4 &b) { return D4(a) * b; }
__device__ __forceinline__ D4 operator/(const D4 &a, double b) { return a / D4(b); }
__device__ __forceinline__ D4 operator/(double a, const D4 &b) { return D4(a) / b; }
__device__ __forceinline__ D4 &operator+=(D4 &a, const D4 &b) { return a = a + b; }
__device__ __forceinline__ D4 &operator-=(D4 &a, const D4 &b) { return a = a - b; }

// a node field whose element m carries the unit derivative in slot `slot`
struct SeedF {
  const double *p;
  int m, slot;
  __device__ __forceinline__ D4 operator[](int k) const {
    D4 r(p[k]);
    if (k == m) r.d[slot] = 1.0;
    return r;
  }
};
// the members the row code reads, with seeded fields
struct SOpD {
  int X, Y;
  const double *gx, *gy;
  SeedF nu, nd, ro, gp;
  double dt, mu;
};

// the right-hand side of a momentum row (k_stokes_rhs): -g rho with g from
// the potential difference
template <class T, class O>
__device__ __forceinline__ T rhs_x(const O &o, int i, int j) {
  const int m = i * o.Y + j;
  const double xkf = (o.gx[i + 1] - o.gx[i - 1]) / 2.0;
  const T GX = -(o.gp[m + o.Y + 1] - o.gp[m + 1]) / T(xkf);
  return -GX * (o.ro[m] + o.ro[m + 1]) / T(2.0);
}
template <class T, class O>
__device__ __forceinline__ T rhs_y(const O &o, int i, int j) {
  const int m = i * o.Y + j;
  const double ykf = (o.gy[j + 1] - o.gy[j - 1]) / 2.0;
  const T GY = -(o.gp[m + o.Y + 1] - o.gp[m + o.Y]) / T(ykf);
  return -GY * (o.ro[m] + o.ro[m + o.Y]) / T(2.0);
}

// ------------------------------------------------------------------ K^T
// Column (comp, m) of K^T y gathers y_r K[r, col] over the rows r of the
// 3x3 node window around m: every row of k_stokes_apply that references an
// unknown of node m sits there (the momentum rows reach one node each way,
// the continuity rows and the continuity-type BC rows, whose cell is shifted
// by at most one, too).  K[r, col] is row r evaluated on the unit vector of
// col.  Ordinary BC rows contribute their diagonal here; their neighbour
// columns, which may be anywhere, come from k_stokes_apply_t_bc.
__global__ void k_stokes_apply_t(SOp<double> o, BcTab bc, const double *__restrict__ y, double *__restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y;
  const int X = o.X, Y = o.Y;
  if (b >= Y) return;
  const int m = a * Y + b;
  double s[3] = {0, 0, 0};
  for (int i = a - 1; i <= a + 1; i++) {
    if (i < 0 || i >= X) continue;
    for (int j = b - 1; j <= b + 1; j++) {
      if (j < 0 || j >= Y) continue;
      const int r = i * Y + j;
      const double y0 = y[3 * r], y1 = y[3 * r + 1], y2 = y[3 * r + 2];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const VUnit<double> e{c, m, Y};
        const bool diag = r == m;
        // continuity / pressure row of node r
        if (y0 != 0) {
          double k;
          if (i && j)
            k = bc.bondm[3 * r] ? (diag && c == 0 ? o.mu : 0.0) : o.mu * cont_g<double>(o, e, i, j);
          else
            k = diag && c == 0 ? o.mu : 0.0;
          s[c] += y0 * k;
        }
        // x- and y-momentum rows of node r
        for (int q = 1; q <= 2; q++) {
          const double yq = q == 1 ? y1 : y2;
          if (yq == 0) continue;
          const bool exists = q == 1 ? j < Y - 1 : i < X - 1;
          double k;
          if (!exists) {
            k = diag && c == q ? o.mu : 0.0;
          } else if (bc.bondm[3 * r + q]) {
            const int bb = bc.bondm[3 * r + q];
            if (bc.bondv[4 * bb + 1] > 1e30) {
              int dm1, dm2;
              cont_shift(bc.bondn[3 * bb] - 1 - (3 * r + q), Y, dm1, dm2);
              k = o.mu * cont_g<double>(o, e, i + dm1, j + dm2);
            } else {
              k = diag && c == q ? o.mu : 0.0;
            }
          } else {
            k = q == 1 ? xmom_g<double>(o, e, i, j, true) : ymom_g<double>(o, e, i, j, true);
          }
          s[c] += yq * k;
        }
      }
    }
  }
  out[3 * m] = s[0];
  out[3 * m + 1] = s[1];
  out[3 * m + 2] = s[2];
}

// the neighbour columns of the ordinary BC rows: row u = mu (x_u - sum k x_nb)
__global__ void k_stokes_apply_t_bc(BcTab bc, int X, int Y, double mu, const double *__restrict__ y,
                                    double *__restrict__ out) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= 3 * X * Y) return;
  const int m = u / 3, q = u % 3, i = m / Y, j = m % Y;
  const bool bcrow = q == 0 ? (i && j) : (q == 1 ? j < Y - 1 : i < X - 1);
  const int b = bc.bondm[u];
  if (!bcrow || !b || y[u] == 0) return;
  if (q != 0 && bc.bondv[4 * b + 1] > 1e30) return;  // continuity-type: gathered
  for (int n = 0; n < 3; n++) {
    const int nb = bc.bondn[3 * b + n];
    if (nb) aadd(&out[nb - 1], -mu * bc.bondv[4 * b + n + 1] * y[u]);
  }
}

// dJ/d(nu, nd, rho, gp)[m] = -sum_r mu_r d(K x - b)_r / d field[m] over the
// momentum rows of the window (the only rows that read node fields)
__global__ void k_stokes_field_adj(SOp<double> o, BcTab bc, const double *__restrict__ x,
                                   const double *__restrict__ lam, double *lnu, double *lnd, double *lro,
                                   double *lgp) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y;
  const int X = o.X, Y = o.Y;
  if (b >= Y) return;
  const int m = a * Y + b;
  const SOpD od{X, Y, o.gx, o.gy, SeedF{o.nu, m, 0}, SeedF{o.nd, m, 1}, SeedF{o.ro, m, 2}, SeedF{o.gp, m, 3},
                o.dt, o.mu};
  const VInter v{x, Y};
  D4 acc(0.0);
  for (int i = a - 1; i <= a + 1; i++) {
    if (i < 0 || i >= X) continue;
    for (int j = b - 1; j <= b + 1; j++) {
      if (j < 0 || j >= Y) continue;
      const int r = i * Y + j;
      if (j < Y - 1 && !bc.bondm[3 * r + 1] && lam[3 * r + 1] != 0)
        acc += lam[3 * r + 1] * (xmom_g<D4>(od, v, i, j, true) - rhs_x<D4>(od, i, j));
      if (i < X - 1 && !bc.bondm[3 * r + 2] && lam[3 * r + 2] != 0)
        acc += lam[3 * r + 2] * (ymom_g<D4>(od, v, i, j, true) - rhs_y<D4>(od, i, j));
    }
  }
  lnu[m] = -acc.d[0];
  lnd[m] = -acc.d[1];
  lro[m] = -acc.d[2];
  lgp[m] = -acc.d[3];
}

// P on the pressure rows (every third unknown): x_p *= f
__global__ void k_scale_p(int N, double f, double *x) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < N) x[3 * m] *= f;
}
__global__ void k_copy_scale_p(int N, double f, const double *a, double *b) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < N) {
    b[3 * m] = a[3 * m] * f;
    b[3 * m + 1] = a[3 * m + 1];
    b[3 * m + 2] = a[3 * m + 2];
  }
}

// The rows the transposed preconditioner eliminates: ordinary BC rows
// mu (x_u - sum k x_nb) and the identity rows of the ghost unknowns (k_stokes_apply);
// continuity-type rows stay with the interior.
__global__ void k_bc_mask(BcTab bc, int X, int Y, int8_t *bm) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= 3 * X * Y) return;
  const int m = u / 3, q = u % 3, i = m / Y, j = m % Y;
  const int b = bc.bondm[u];
  bool B;
  if (q == 0)
    B = !(i && j) || b;
  else if (q == 1 ? j == Y - 1 : i == X - 1)
    B = true;
  else
    B = b && !(bc.bondv[4 * b + 1] > 1e30);
  bm[u] = B ? 1 : 0;
}
__global__ void k_mask_keep(int n, const int8_t *bm, int keep, const double *a, double *b) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u < n) b[u] = bm[u] == keep ? a[u] : 0.0;
}
// r' = r - C_IB r_B / mu on the interior rows, 0 on the eliminated ones
__global__ void k_fold(int n, const int8_t *bm, double mu, const double *r, const double *t, double *o) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u < n) o[u] = bm[u] ? 0.0 : r[u] - t[u] / mu;
}
// z_B = (r_B - C_BI z_I) / mu
__global__ void k_unfold(int n, const int8_t *bm, double mu, const double *r, const double *t, double *z) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u < n && bm[u]) z[u] = (r[u] - t[u]) / mu;
}

// the markers of the objective (inside, markk > 0, active type, iron, mobile)
__device__ __forceinline__ bool sink_marker(const i2e_params *p, const Grid &g, float fx, float fy, float fk, int t,
                                            double &m, double &rx, double &ry) {
  const int mm2 = t >= 100 ? t - 100 : t;
  if (!(fx > 0 && fy > 0 && (double)fx < g.xsize && (double)fy < g.ysize && fk > 0 && t < 50)) return false;
  if ((unsigned)mm2 >= (unsigned)I2E_MAXTYPES || !p->is_iron[mm2] || p->immobile[mm2]) return false;
  const double dx = (double)fx - g.xsize / 2.0, dy = (double)fy - g.ysize / 2.0, r = sqrt(dx * dx + dy * dy);
  if (!(r > 0)) return false;
  m = p->rock[mm2].ro;
  rx = dx / r;
  ry = dy / r;
  return true;
}

__global__ __launch_bounds__(RB) void k_sink_part(const i2e_params *__restrict__ p, Grid g, NodeF F, Markers mk,
                                                  int64_t M, double dt, double *part) {
  double smv = 0, sm = 0;
  for (int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x; i < M; i += (int64_t)RG * RB) {
    double m, rx, ry;
    if (sink_marker(p, g, mk.x[i], mk.y[i], mk.k[i], mk.t[i], m, rx, ry)) {
      Rk4 R;
      rk4(p, g, F.vx, F.vy, (double)mk.x[i], (double)mk.y[i], dt, R);
      smv += m * (R.vx * rx + R.vy * ry);
      sm += m;
    }
  }
  // block sums in a fixed order
  __shared__ double sa[RB / 64], sb[RB / 64];
  smv = wave_sum(smv);
  sm = wave_sum(sm);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sa[w] = smv;
    sb[w] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int k = 0; k < RB / 64; k++) {
      ta += sa[k];
      tb += sb[k];
    }
    part[2 * blockIdx.x] = ta;
    part[2 * blockIdx.x + 1] = tb;
  }
}

// reverse of the RK4 velocity of every objective marker: dJ/d(node vx, vy)
// into g (the Stokes vector layout, 3 m + 1 and 3 m + 2)
__global__ __launch_bounds__(256) void k_sink_adj(const i2e_params *__restrict__ p, Grid g, NodeF F, Markers mk,
                                                  int64_t M, double dt, double msum, double *gv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    double m, rx, ry;
    if (!sink_marker(p, g, mk.x[i], mk.y[i], mk.k[i], mk.t[i], m, rx, ry)) continue;
    Rk4 R;
    rk4(p, g, F.vx, F.vy, (double)mk.x[i], (double)mk.y[i], dt, R);
    const double gx = m / msum * rx, gy = m / msum * ry;
    const int Y = g.Y;
    if (p->markmod == 1) {
      scatter_w(R.s[0].wx, Y, gx, gv, 1);
      scatter_w(R.s[0].wy, Y, gy, gv, 2);
      continue;
    }
    // stage velocity adjoints; stage k+1's position is x0 + v_k h_k
    double bx[4] = {gx / 6.0, gx / 3.0, gx / 3.0, gx / 6.0}, by[4] = {gy / 6.0, gy / 3.0, gy / 3.0, gy / 6.0};
    const double h[4] = {0.0, dt / 2.0, dt / 2.0, dt};
    for (int k = 3; k >= 0; k--) {
      const Interp &I = R.s[k];
      scatter_w(I.wx, Y, bx[k], gv, 1);
      scatter_w(I.wy, Y, by[k], gv, 2);
      if (k == 0) break;
      double px = 0, py = 0;
      interp_grad_t(g, I.wx, 0, +1, F.vx, bx[k], I.cx, I.cy, px, py);
      interp_grad_t(g, I.wy, +1, 0, F.vy, by[k], I.cx, I.cy, px, py);
      bx[k - 1] += px * h[k];
      by[k - 1] += py * h[k];
    }
  }
}

// ------------------------------------------------------------------ viscosity scatter
// the marker's viscosity as ronurecalc takes it (the stored markv when
// positive, else viscalc) and d eta / d s, s scaling every markdh
__device__ __forceinline__ bool visc_marker(const i2e_params *p, const Grid &g, const NodeF &F, const Markers &mk,
                                            int64_t i, double &x, double &y, double &eta, double &deta) {
  const int t = mk.t[i];
  const int mm2 = t >= 100 ? t - 100 : t;
  const float fx = mk.x[i], fy = mk.y[i], fk = mk.k[i];
  if (!(fx > 0 && fy > 0 && (double)fx < g.xsize && (double)fy < g.ysize && fk > 0 && t < 50)) return false;
  x = fx;
  y = fy;
  double mtk = (double)fk;
  if (mm2 < 2) mtk = (double)(float)p->tmp_ambient;
  const int ty = t < 20 ? mm2 : mm2 - 20;
  const double mpb = 0.0;  // viscalc reads P only through markdv, refused here
  const double raw = viscalc_d(p, mtk, mpb, ty, 0.0);
  eta = mk.v[i] > 0 ? (double)mk.v[i] : raw;
  deta = 0;
  const i2e_rock &r = p->rock[ty];
  if (ty <= 20 && r.nu0 != 0 && r.dh > 0 && (r.ss == 0 || r.mm == 1.0)) {
    const double e1 = (r.dh + r.dv * mpb) / (8.314 * mtk);
    const double unc = r.nu0 * exp(e1);  // viscalc before its limits
    const bool lim = e1 > 150.0 || unc <= r.n0 || unc >= r.n1 || unc <= p->nubeg || unc >= p->nuend;
    if (!lim) deta = raw * r.dh / (8.314 * mtk);
  }
  const double cx = (double)fx - g.xsize / 2.0, cy = (double)fy - g.ysize / 2.0;
  if (sqrt(cx * cx + cy * cy) > g.xsize / 2.0 * p->radmax_frac) {
    eta = p->eta_outer;
    deta = 0;
  }
  return true;
}

__global__ void k_visc_sums(const i2e_params *__restrict__ p, Grid g, NodeF F, Markers mk, int64_t M, double *S) {
  const int N = g.X * g.Y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    double x, y, eta, deta;
    if (!visc_marker(p, g, F, mk, i, x, y, eta, deta)) continue;
    const Tent T = tent(g, x, y);
    if (T.nn >= 0 && T.wn != 0) {
      aadd(&S[T.nn], eta * T.wn);
      aadd(&S[N + T.nn], T.wn);
    }
    if (T.nc >= 0 && T.wc != 0) {
      aadd(&S[2 * N + T.nc], eta * T.wc);
      aadd(&S[3 * N + T.nc], T.wc);
    }
  }
}

// node normalisation and clamp transposed: l <- l / W where the node took
// the weighted mean unclamped, 0 elsewhere.  A node value that is neither
// the mean nor its [nubeg, nuend] clamp was set by the contrast limit, which
// couples every node to the smallest: flagged, not differentiated.
__global__ void k_visc_nodes_adj(int N, double nubeg, double nuend, const double *nu, const double *nd,
                                 const double *S, double *lnu, double *lnd, int *flag) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  for (int f = 0; f < 2; f++) {
    const double W = S[(2 * f + 1) * N + m];
    double *l = f ? lnd : lnu;
    if (W == 0) {
      l[m] = 0;  // no marker: the node kept its previous value
      continue;
    }
    const double v = S[2 * f * N + m] / W, vc = v < nubeg ? nubeg : (v > nuend ? nuend : v);
    const double have = f ? nd[m] : nu[m];
    if (dabs(have - vc) > 1e-9 * vc) atomicOr(flag, 1);
    l[m] = v == vc ? l[m] / W : 0.0;
  }
}

// per marker: dJ/ds_i = (sum over its tent nodes of l w) d eta / d s
__global__ void k_visc_adj(const i2e_params *__restrict__ p, Grid g, NodeF F, Markers mk, int64_t M,
                           const double *lnu, const double *lnd, double *gs) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    double x, y, eta, deta;
    double r = 0;
    if (visc_marker(p, g, F, mk, i, x, y, eta, deta) && deta != 0) {
      const Tent T = tent(g, x, y);
      double le = 0;
      if (T.nn >= 0) le += lnu[T.nn] * T.wn;
      if (T.nc >= 0) le += lnd[T.nc] * T.wc;
      r = le * deta;
    }
    gs[i] = r;
  }
}

}  // namespace i2e

using namespace i2e;

static inline dim3 nb2f(int X, int Y) { return dim3((Y + 255) / 256, X); }
static inline unsigned mgridf(int64_t M) { return nblk(M, 256) < 65535u ? nblk(M, 256) : 65535u; }

// Rheologies whose viscosity derivative in s is built: Newtonian Arrhenius
// (or constant) without the activation-volume and negative-dh branches, no
// melting switch.
static int flow_adj_supported(const i2e_ctx *c) {
  const i2e_params &P = c->P;
  if (P.melt_on) {
    set_error("flow adjoint: the melting viscosity switch is not differentiated");
    return I2E_ERR_UNSUPPORTED;
  }
  for (int t = 0; t < I2E_MAXTYPES; t++) {
    const i2e_rock &r = P.rock[t];
    if (r.nu0 != 0 && (r.dv != 0 || r.dh < 0 || (r.ss != 0 && r.mm != 1.0))) {
      set_error("flow adjoint: rock type %d has an activation volume, a negative activation energy or a power "
                "law, which are not differentiated", t);
      return I2E_ERR_UNSUPPORTED;
    }
  }
  return I2E_OK;
}

static int sink_sums(i2e_ctx *c, double *smv, double *sm) {
  if (c->fadj.n < 2 * RG + 2) ST_OK(c->fadj.alloc(2 * RG + 2));
  double *part = c->fadj.p;
  hipLaunchKernelGGL(k_sink_part, dim3(RG), dim3(RB), 0, c->s, c->dP, c->grid(), c->F, c->mk, c->M, c->timestep,
                     part);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RB), 0, c->s, part, part + 2 * RG);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(c->pin.p, part + 2 * RG, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  *smv = c->pin.p[0];
  *sm = c->pin.p[1];
  return I2E_OK;
}

// K^T mu = g for the Stokes rows of the context's current fields and
// set-up (stokes_dt the stabilisation's dt); g is overwritten.  scratch:
// three vectors of 3N.
int i2e::stokes_solve_t(i2e_ctx *c, double *gv, double *mu, double *scratch, int *its_out, double *rr_out) {
  const int X = c->X, Y = c->Y, N = c->N;
  const size_t n3 = 3 * (size_t)N;
  ST_OK(c->fmask.alloc(n3));
  int8_t *bm = c->fmask.p;
  double *tmp = scratch, *tb = scratch + n3, *tc = scratch + 2 * n3;
  hipLaunchKernelGGL(k_bc_mask, dim3(nblk((long)n3, 256)), dim3(256), 0, c->s, c->bc, X, Y, bm);
  HIP_OK(hipMemsetAsync(mu, 0, sizeof(double) * n3, c->s));
  SOp<double> o = c->sop();
  o.dt = c->stokes_dt;
  // K^T mu = g, solved as (P^-1 K^T P) nu = P^-1 g, mu = P nu: where the
  // viscous block is symmetric P^-1 K^T P = K, so the forward preconditioner
  // applies unchanged and the residual is measured with the forward rows'
  // scaling (continuity rows times mukoef)
  auto A = [&](const double *x, double *y) -> int {
    hipLaunchKernelGGL(k_copy_scale_p, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, 1.0 / o.mu, x, tmp);
    hipLaunchKernelGGL(k_stokes_apply_t, nb2f(X, Y), dim3(256), 0, c->s, o, c->bc, (const double *)tmp, y);
    hipLaunchKernelGGL(k_stokes_apply_t_bc, dim3(nblk((long)n3, 256)), dim3(256), 0, c->s, c->bc, X, Y, o.mu,
                       (const double *)tmp, y);
    hipLaunchKernelGGL(k_scale_p, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, o.mu, y);
    return I2E_OK;
  };
  // preconditioner of C = P^-1 K^T P: eliminate the BC rows, whose
  // transposes couple a ghost unknown to the interior rows with weight mu
  // (C_BB = mu I), apply the forward preconditioner to the interior
  // (reduced operator C_II - C_IB C_BI / mu, the transpose of the reduced
  // forward operator up to P), and back-substitute:
  //   r'_I = r_I - C_IB r_B / mu,  z_I = M r'_I,  z_B = (r_B - C_BI z_I) / mu
  const unsigned gb = nblk((long)n3, 256);
  auto Mf = [&](const double *r, double *z) -> int {
    hipLaunchKernelGGL(k_mask_keep, dim3(gb), dim3(256), 0, c->s, (int)n3, bm, 1, r, tb);
    ST_OK(A(tb, tc));
    hipLaunchKernelGGL(k_fold, dim3(gb), dim3(256), 0, c->s, (int)n3, bm, o.mu, r, (const double *)tc, tb);
    ST_OK(stokes_precond(c, tb, z, false));
    hipLaunchKernelGGL(k_mask_keep, dim3(gb), dim3(256), 0, c->s, (int)n3, bm, 0, (const double *)z, tb);
    ST_OK(A(tb, tc));
    hipLaunchKernelGGL(k_unfold, dim3(gb), dim3(256), 0, c->s, (int)n3, bm, o.mu, r, (const double *)tc, z);
    return I2E_OK;
  };
  hipLaunchKernelGGL(k_scale_p, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, o.mu, gv);
  c->fs.sx = c->fs.sy = nullptr;
  c->fs.split_of = nullptr;
  int its = 0;
  double rr = 0;
  int st = c->fs.solve(A, Mf, gv, mu, c->o.stokes_rtol, c->o.maxit, c->s, &its, &rr);
  *its_out = its;
  *rr_out = rr;
  if (st) {
    set_error("adjoint: transposed Stokes solve, %d iterations, relative residual %.3e > %.1e", its, rr,
              c->o.stokes_rtol);
    return st;
  }
  hipLaunchKernelGGL(k_scale_p, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, 1.0 / o.mu, mu);
  return I2E_OK;
}

extern "C" {

int i2e_stokes_apply_t_dev(i2e_ctx *c, const double *x, double *y) {
  if (!c || !x || !y || !c->setup_done) return I2E_ERR_ARG;
  SOp<double> o = c->sop();
  hipLaunchKernelGGL(k_stokes_apply_t, nb2f(c->X, c->Y), dim3(256), 0, c->s, o, c->bc, x, y);
  hipLaunchKernelGGL(k_stokes_apply_t_bc, dim3(nblk(3 * (long)c->N, 256)), dim3(256), 0, c->s, c->bc, c->X, c->Y,
                     o.mu, x, y);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->s));
  return I2E_OK;
}

int i2e_sink_rate(i2e_ctx *c, double *J) {
  if (!c || !J || !c->setup_done) return I2E_ERR_ARG;
  double smv, sm;
  ST_OK(sink_sums(c, &smv, &sm));
  *J = sm > 0 ? smv / sm : 0.0;
  return I2E_OK;
}

int i2e_adjoint_sink_rate(i2e_ctx *c, i2e_flow_adjoint_out *out) {
  if (!c || !out || !c->setup_done) return I2E_ERR_ARG;
  if (!c->stokes_valid) {
    set_error("flow adjoint: no Stokes solution of the current markers (call viterate, and nothing that changes "
              "the markers or fields, first)");
    return I2E_ERR_ARG;
  }
  ST_OK(flow_adj_supported(c));
  memset(out, 0, sizeof *out);
  const int X = c->X, Y = c->Y, N = c->N;
  const int64_t M = c->M;
  const Grid g = c->grid();
  double smv, sm;
  ST_OK(sink_sums(c, &smv, &sm));
  if (!(sm > 0)) {
    set_error("flow adjoint: no iron marker inside the grid");
    return I2E_ERR_ARG;
  }
  out->J = smv / sm;
  // device scratch, kept by the context: the reduction parts, g, mu, the
  // solution x and three operator/preconditioner temporaries (3N each),
  // the four field adjoints and the four viscosity sums (N each), the
  // per-marker gradient; the eliminated-row mask (3N bytes)
  const size_t n3 = 3 * (size_t)N;
  ST_OK(c->fadj.alloc(2 * RG + 2 + 6 * n3 + 8 * (size_t)N + (M ? M : 1)));
  ST_OK(c->fflag.alloc(1));
  double *part = c->fadj.p, *gv = part + 2 * RG + 2, *mu = gv + n3, *xs = mu + n3, *tmp = xs + n3,
         *lnu = tmp + 3 * n3, *lnd = lnu + N, *lro = lnd + N, *lgp = lro + N, *S = lgp + N, *gs = S + 4 * (size_t)N;
  DBuf<int> &flag = c->fflag;
  hipLaunchKernelGGL(k_stokes_pack, nb2f(X, Y), dim3(256), 0, c->s, X, Y, c->F.pr, c->F.vx, c->F.vy, xs);
  HIP_OK(hipMemsetAsync(gv, 0, sizeof(double) * n3, c->s));
  HIP_OK(hipMemsetAsync(S, 0, sizeof(double) * 4 * (size_t)N, c->s));
  HIP_OK(hipMemsetAsync(flag.p, 0, sizeof(int), c->s));
  const unsigned gm = mgridf(M ? M : 1);
  if (M) hipLaunchKernelGGL(k_sink_adj, dim3(gm), dim3(256), 0, c->s, c->dP, g, c->F, c->mk, M, c->timestep, sm, gv);
  int its = 0;
  double rr = 0;
  const int st = stokes_solve_t(c, gv, mu, tmp, &its, &rr);
  out->stokes_iters = its;
  out->stokes_relres = rr;
  if (st) return st;
  SOp<double> o = c->sop();
  o.dt = c->stokes_dt;
  hipLaunchKernelGGL(k_stokes_field_adj, nb2f(X, Y), dim3(256), 0, c->s, o, c->bc, (const double *)xs,
                     (const double *)mu, lnu, lnd, lro, lgp);
  if (M) hipLaunchKernelGGL(k_visc_sums, dim3(gm), dim3(256), 0, c->s, c->dP, g, c->F, c->mk, M, S);
  hipLaunchKernelGGL(k_visc_nodes_adj, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, c->P.nubeg, c->P.nuend, c->F.nu,
                     c->F.nd, S, lnu, lnd, flag.p);
  if (M) hipLaunchKernelGGL(k_visc_adj, dim3(gm), dim3(256), 0, c->s, c->dP, g, c->F, c->mk, M, lnu, lnd, gs);
  hipLaunchKernelGGL(k_sum_part, dim3(RG), dim3(RB), 0, c->s, (int64_t)M, gs, gs, part);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RB), 0, c->s, part, part + 2 * RG);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(c->pin.p, part + 2 * RG, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipMemcpyAsync(c->pin.p + 2, flag.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  int fl;
  memcpy(&fl, c->pin.p + 2, sizeof fl);
  if (fl) {
    set_error("flow adjoint: the viscosity contrast limit (nucontr) set some node viscosities; not differentiated");
    return I2E_ERR_UNSUPPORTED;
  }
  out->dJ_dEscale = c->pin.p[0];
  return I2E_OK;
}

}  // extern "C"
