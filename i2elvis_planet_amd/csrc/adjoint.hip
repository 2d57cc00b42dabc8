// adjoint.hip -- discrete adjoint of a peak-temperature objective along a
// recorded trajectory, and the transposed heat operator.
//
// Objective (a sensitivity study of the planetesimal runs; no counterpart
// in the reference):
//   J = sum_i m_i marktmax_i / sum_i m_i
// over silicate markers (is_silicate) inside the grid, with the mass of an
// equal-volume marker m_i = markro (1 - markpor_i).  Parameters: h0_scale
// (the multiplier of the 26Al heating, markmart.c:758-770) and phi0, a
// uniform shift of every marker porosity that is positive at the start.
//
// The model differentiated is the thermal chain of the time step
// (i2mart.c:60-153) with marker positions and time steps held at their
// recorded values:
//   ronurecalc: property and temperature scatter, node normalisation and
//               tkrecalc's BC rows (markmart.c:192-1019, heatmart.c:758-826),
//               marktmax update (markmart.c:354-360)
//   titerate:   implicit heat rows and solve (heatmart.c:406-439), marker
//               temperature update with subgrid diffusion and its node
//               correction (heatmart.c:515-689)
// Stokes, gravity and advection enter only through the recorded positions
// and time steps; shear heating enters as the recorded source.  The
// gradient is therefore the full-model gradient whenever the flow does not
// depend on the parameters (temperature-independent density, viscosities at
// their clip limits); otherwise it is the frozen-trajectory gradient, which
// i2e_replay evaluates exactly.
//
// The whole history (positions and temperatures of every marker after each
// ronurecalc, six node fields per step) stays resident in device memory:
// 12 B per marker per step plus 48 B per node per step.
#include <string.h>

#include "adj_dev.hpp"
#include "viscalc.hpp"

namespace i2e {

struct AdjHist {
  int K = 0;
  int64_t M = 0;
  bool valid = false, shear = false;
  int limited = 0;           // steps whose dt titerate's heating-rate limit shrank
  DBuf<float> mx, my, mT;    // (K+1) x M, after ronurecalc r
  DBuf<int16_t> last;        // ronurecalc index of the last marktmax increase, -1 none
  DBuf<float> tmax0;         // marktmax before a ronurecalc
  DBuf<double> nodes;        // K x 6 x N: tk_old, tk_new, ro, cp, kt, sol1
  std::vector<double> dt, tsum;
  std::vector<int> solved;   // titerate ran its solve in step n
  // flow part of the record (the coupled adjoint): per step the fields the
  // Stokes rows read and their solution, the viscosity each ronurecalc read
  DBuf<double> fnodes;       // K x 7 x N: nu, nd, gp, ht, vx, vy, pr after viterate n
  DBuf<float> meta;          // (K+1) x M: markv as ronurecalc r read it
  std::vector<double> sdt;   // stabilisation dt of the Stokes solve of step n
  int vlimited = 0;          // steps whose dt maxvelstep set below maxtmstep
  // reverse pass
  DBuf<double> lmax, lamT, lamTp, lpor, lh0;  // per marker
  DBuf<double> L;            // 2 x NADJ x N node adjoints
  DBuf<double> nw;           // W, S0, LS, mu, nu, tmp (6 x N)
  DBuf<double> part;
  DBuf<double> lX, leta, ls; // coupled: position (2M), viscosity and per-marker dJ/ds adjoints
  DBuf<double> fw;           // coupled: flow work vectors
  bool lx_valid = false;     // lX holds dJ/d(initial positions) of this record
};

void adj_release(AdjHist *h) {
  if (!h) return;
  h->mx.release();
  h->my.release();
  h->mT.release();
  h->last.release();
  h->tmax0.release();
  h->nodes.release();
  h->lmax.release();
  h->lamT.release();
  h->lamTp.release();
  h->lpor.release();
  h->lh0.release();
  h->L.release();
  h->nw.release();
  h->part.release();
  h->fnodes.release();
  h->meta.release();
  h->lX.release();
  h->leta.release();
  h->ls.release();
  h->fw.release();
  delete h;
}

// node adjoints of the fields one ronurecalc + titerate (+ viterate) pass
// reads/writes: NADJ arrays of N
constexpr int NADJ = 8;
struct NodeAdj {
  double *tk, *ro, *cp, *kt, *ht, *tnew, *nu, *nd;
};
static NodeAdj node_adj(double *base, int N) {
  return NodeAdj{base, base + N, base + 2 * (size_t)N, base + 3 * (size_t)N, base + 4 * (size_t)N,
                 base + 5 * (size_t)N, base + 6 * (size_t)N, base + 7 * (size_t)N};
}

__device__ __forceinline__ bool inside_strict(const Grid &g, float fx, float fy) {
  return fx > 0 && fy > 0 && (double)fx < g.xsize && (double)fy < g.ysize;
}
// movemark's out-of-grid test (markk = 0 afterwards), markmart.c:149-152
__device__ __forceinline__ bool outside(const Grid &g, float fx, float fy) {
  return fx < 0 || fy < 0 || (double)fx > g.xsize || (double)fy > g.ysize;
}

// Wave-segmented scatter of NV four-corner values per lane (lanes of a
// wavefront in the same cell are summed before one atomic per corner).
template <int NV>
struct Dst {
  double *p[NV];
};
template <int NV>
__device__ __forceinline__ void seg_scatter(int key, double (&v)[NV][4], const Dst<NV> &d, int Y, int lane) {
  const SegScan sc(key, lane);
#pragma unroll
  for (int q = 0; q < NV; q++)
#pragma unroll
    for (int c = 0; c < 4; c++) v[q][c] = sc.sum(v[q][c]);
  if (sc.tail && key >= 0) {
    const int m10 = key / Y, m20 = key % Y;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int m = (m10 + (c >> 1)) * Y + m20 + (c & 1);
#pragma unroll
      for (int q = 0; q < NV; q++)
        if (v[q][c] != 0) aadd(&d.p[q][m], v[q][c]);
    }
  }
}

__device__ __forceinline__ double corner_w(const W4 &w, int c) {
  return ((c >> 1) ? w.wx1 : w.wx0) * ((c & 1) ? w.wy1 : w.wy0);
}

// ---------------------------------------------------------------- forward record
__global__ void k_last_set(int64_t M, const float *before, const float *after, int r, int16_t *last) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x)
    if (after[i] != before[i]) last[i] = (int16_t)r;
}

// positions of the recorded step; movemark's out-of-grid rule on markk
__global__ void k_replay_move(Grid g, Markers mk, int64_t M, const float *x, const float *y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const float fx = x[i], fy = y[i];
    mk.x[i] = fx;
    mk.y[i] = fy;
    if (outside(g, fx, fy)) mk.k[i] = 0;
  }
}

// ---------------------------------------------------------------- objective
__device__ __forceinline__ bool objective_marker(const i2e_params *p, const Grid &g, float fx, float fy, float fk,
                                                 int t, float por, double &m) {
  const int mm2 = t >= 100 ? t - 100 : t;
  if (!(inside_strict(g, fx, fy) && fk > 0 && t < 50 && p->is_silicate[mm2])) return false;
  double pe = (mm2 == 0 || mm2 >= 7) ? 0.0 : (por < 0.0f ? 0.0 : (por > 1.0f ? 1.0 : (double)por));
  m = p->rock[mm2].ro * (1.0 - pe);
  return true;
}


__device__ __forceinline__ void block_store2(double a, double b, double *out) {
  __shared__ double sa[RB / 64], sb[RB / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sa[w] = a;
    sb[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int k = 0; k < RB / 64; k++) {
      ta += sa[k];
      tb += sb[k];
    }
    out[2 * blockIdx.x] = ta;
    out[2 * blockIdx.x + 1] = tb;
  }
}

__global__ __launch_bounds__(RB) void k_peak_part(const i2e_params *__restrict__ p, Grid g, Markers mk, int64_t M,
                                                  double *part) {
  double smt = 0, sm = 0;
  for (int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x; i < M; i += (int64_t)RG * RB) {
    double m;
    if (objective_marker(p, g, mk.x[i], mk.y[i], mk.k[i], mk.t[i], mk.por[i], m)) {
      smt += m * (double)mk.tmax[i];
      sm += m;
    }
  }
  block_store2(smt, sm, part);
}

__global__ __launch_bounds__(RB) void k_sum_part(int64_t n, const double *a, const double *b, double *part) {
  double sa = 0, sb = 0;
  for (int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x; i < n; i += (int64_t)RG * RB) {
    sa += a[i];
    sb += b[i];
  }
  block_store2(sa, sb, part);
}

__global__ __launch_bounds__(RB) void k_sum_final(const double *part, double *out) {
  double a = 0, b = 0;
  for (int k = threadIdx.x; k < RG; k += RB) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  __shared__ double o[2];
  block_store2(a, b, o);
  if (threadIdx.x == 0) {
    out[0] = o[0];
    out[1] = o[1];
  }
}

// dJ/dmarktmax_i = m_i / sum m; dJ/dpor_i through m_i
__global__ void k_adj_init(const i2e_params *__restrict__ p, Grid g, Markers mk, int64_t M, double J, double msum,
                           double *lmax, double *lpor) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    double m;
    const float por = mk.por[i];
    const int t = mk.t[i];
    if (objective_marker(p, g, mk.x[i], mk.y[i], mk.k[i], t, por, m)) {
      const int mm2 = t >= 100 ? t - 100 : t;
      lmax[i] = m / msum;
      const bool live = !(mm2 == 0 || mm2 >= 7) && por > 0.0f && por < 1.0f;
      lpor[i] = live ? -p->rock[mm2].ro * ((double)mk.tmax[i] - J) / msum : 0.0;
    } else {
      lmax[i] = 0;
      lpor[i] = 0;
    }
  }
}

// ---------------------------------------------------------------- reverse: ronurecalc
struct Slot {
  const float *x, *y, *T;
};

// W = sum of the scatter weights of the markers active in ronurecalc
// (in the grid, markk > 0, type < 50: markmart.c:231-236)
__global__ __launch_bounds__(256) void k_adj_weights(Grid g, Slot s, const int8_t *mt, int64_t M, double *W) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < M; base += stride) {
    const int64_t i = base + threadIdx.x;
    double v[1][4] = {{0, 0, 0, 0}};
    int key = -1;
    if (i < M) {
      const float fx = s.x[i], fy = s.y[i];
      if (inside_strict(g, fx, fy) && s.T[i] > 0 && mt[i] < 50) {
        const double x = fx, y = fy;
        const int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
        W4 w;
        nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
        key = m10 * g.Y + m20;
#pragma unroll
        for (int c = 0; c < 4; c++) v[0][c] = corner_w(w, c);
      }
    }
    seg_scatter<1>(key, v, Dst<1>{{W}}, g.Y, lane);
  }
}

// tkrecalc rows T_m = c + sum k T_nbr (k_tbc) transposed: the BC node's
// adjoint moves to its neighbours and is cleared
__global__ void k_adj_tbc(BcTab bc, int N, double *ltk) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  int b = bc.bondm[3 * N + m];
  if (!b) return;
  const double v = ltk[m];
  ltk[m] = 0;
  for (int n1 = 0; n1 < 3; n1++) {
    int nb = bc.bondn[3 * b + n1];
    if (nb && v != 0) aadd(&ltk[nb - 1 - 3 * N], bc.bondv[4 * b + n1 + 1] * v);
  }
}

// node normalisation (k_ronu_nodes) transposed: covered nodes divide by the
// weight sum; uncovered nodes keep their previous values, so their adjoints
// pass to the previous step's fields (node T: titerate's solution)
__global__ void k_adj_ronu_nodes(int N, const double *W, NodeAdj L, NodeAdj Lp, int has_prev) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  const double w = W[m];
  if (w != 0) {
    L.tk[m] /= w;
    L.ro[m] /= w;
    L.cp[m] /= w;
    L.kt[m] /= w;
    L.ht[m] /= w;
  } else {
    if (has_prev) {
      Lp.tnew[m] += L.tk[m];
      Lp.ro[m] += L.ro[m];
      Lp.cp[m] += L.cp[m];
      Lp.kt[m] += L.kt[m];
      Lp.ht[m] += L.ht[m];
    }
    L.tk[m] = L.ro[m] = L.cp[m] = L.kt[m] = L.ht[m] = 0;
  }
}


// porosity derivative of the conductivity closure of porous types
// (markmart.c:614-638, as k_ronu_markers); 0 where a clamp holds
__device__ __forceinline__ double dkt_dpor(const i2e_params *p, const i2e_rock &r, float por) {
  const double kt0 = r.kt, pr0 = (double)por;
  double k, dk;
  if (por < 0.200f) {
    k = kt0 * exp(-pr0 / 0.080);
    dk = -k / 0.080;
  } else if (por <= 0.400f) {
    const double a = kt0 * exp(-pr0 / 0.080), b = kt0 * exp(-1.200 - (pr0 / 0.167));
    const double a4 = pow(a, 4.000), b4 = pow(b, 4.000);
    k = pow(a4 + b4, 0.250);
    dk = pow(a4 + b4, -0.750) * (a4 * (-1.0 / 0.080) + b4 * (-1.0 / 0.167));
  } else {
    k = kt0 * exp(-1.200 - (pr0 / 0.167));
    dk = -k / 0.167;
  }
  if (k <= 0.0010 || k >= p->k_cutoff) return 0.0;
  return dk;
}

// ronurecalc's marker pass transposed (markmart.c:281-360, 475-785, 840-911):
// lamT (adjoint of markk after the pass) -> lamTp (before the pass), plus the
// marktmax, density, conductivity and radiogenic-heat paths into the
// per-marker porosity and h0_scale gradients.
__global__ void k_adj_ronu_markers(const i2e_params *__restrict__ p, Grid g, Slot s, const int8_t *mt,
                                   const float *por_a, int64_t M, const int16_t *last, int r, const double *lmax,
                                   const double *lamT, double *lamTp, double *lpor, double *lh0, NodeAdj L,
                                   double timesum, int moved) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const float fx = s.x[i], fy = s.y[i];
    const int t = mt[i];
    if (!(inside_strict(g, fx, fy) && s.T[i] > 0 && t < 50)) {
      // untouched by the pass; movemark zeroed markk of markers that left
      lamTp[i] = (moved && outside(g, fx, fy)) ? 0.0 : lamT[i];
      continue;
    }
    const int mm2 = t >= 100 ? t - 100 : t;
    const double x = fx, y = fy;
    const int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
    W4 w;
    nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
    const double gT = isum(w, g.Y, L.tk), gro = isum(w, g.Y, L.ro), gkt = isum(w, g.Y, L.kt),
                 ght = isum(w, g.Y, L.ht);
    double lt = 0;
    if (mm2 >= 2) lt = lamT[i] + gT + ((t != 0 && last[i] == r) ? lmax[i] : 0.0);
    lamTp[i] = lt;
    const i2e_rock &rk = p->rock[mm2];
    const float por = por_a[i];
    const bool live = !(mm2 == 0 || mm2 >= 7) && por > 0.0f && por < 1.0f;
    const double pe = (mm2 == 0 || mm2 >= 7) ? 0.0 : (por < 0.0f ? 0.0 : (por > 1.0f ? 1.0 : (double)por));
    const double mro = p->densimod == 0 ? rk.ro : rk.ro * (1.000 - pe);
    const double timesumA = timesum / (365.250 * 24.000 * 3600.000);
    double hsrc = 0, h0part = 0;  // mht = hsrc * mro
    if (p->is_silicate[mm2]) {
      const double init_al = 1.89716e+23 * p->al2627_init * 4.999e-13 / (3.264e+13);
      double ll = 0;
      if (p->longlived != 0)
        ll = 0.40689 * 1.430e-11 * exp(-timesumA * 4.9867e-10) + 2.990e-12 * exp(-timesumA * 9.8458e-10) +
             1.600e-12 * exp(-timesumA * 1.5541e-10) + 1.000e-12 * exp(-timesumA * 4.9511e-11);
      h0part = init_al * exp(-timesumA * 9.6673e-7);
      hsrc = h0part * p->h0_scale + ll;
    } else if (p->is_iron[mm2]) {
      const double init_fe = 1.971e+24 * p->fe6056_init * 4.3451e-13 / (1.183e+14);
      double ka = 0;
      if (p->longlived != 0) ka = 0.090625 * 1.430e-11 * exp(-timesumA * 4.9867e-10);
      hsrc = init_fe * exp(-timesumA * 2.6660e-7) + ka;
    }
    lh0[i] += ght * h0part * mro;
    if (live) {
      const double dmro = p->densimod == 0 ? 0.0 : -rk.ro;
      const double dmkt = p->porous[mm2] ? dkt_dpor(p, rk, por) : 0.0;
      lpor[i] += gro * dmro + ght * hsrc * dmro + gkt * dmkt;
    }
  }
}

// ---------------------------------------------------------------- reverse: titerate markers
// pass 1: S0 = sum w over the markers titerate updates (heatmart.c:529-537),
// LS = -sum w lamT' (the antidiffusion correction's adjoint)
__global__ __launch_bounds__(256) void k_adj_heat_mk1(Grid g, Slot s, const int8_t *mt, int64_t M,
                                                      const double *lamTp, double *S0, double *LS) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < M; base += stride) {
    const int64_t i = base + threadIdx.x;
    double v[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    int key = -1;
    if (i < M) {
      const float fx = s.x[i], fy = s.y[i];
      if (inside_strict(g, fx, fy) && mt[i] < 50) {
        const double x = fx, y = fy;
        const int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
        W4 w;
        nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
        key = m10 * g.Y + m20;
        const double l = lamTp[i];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          v[0][c] = corner_w(w, c);
          v[1][c] = -l * v[0][c];
        }
      }
    }
    seg_scatter<2>(key, v, Dst<2>{{S0, LS}}, g.Y, lane);
  }
}

__global__ void k_adj_div(int N, const double *S0, double *LS) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < N) LS[m] = S0[m] != 0 ? LS[m] / S0[m] : 0.0;
}

// step-n node fields of the history
struct StepF {
  const double *tko, *tkn, *ro, *cp, *kt, *sol1;
};

// pass 2: per marker, the transposed update
//   T' = T + dtk1 + (e2 - e3) - sum_n w_n S1_n / S0_n,
//   dtk1 = (e3 - T)(1 - exp(-D)),  D = heatdif kt/(ro cp) 2(1/dx^2+1/dy^2) dt
// with e2, e3 and kt, ro, cp interpolated from the nodes (heatmart.c:556-627)
__global__ __launch_bounds__(256) void k_adj_heat_mk2(const i2e_params *__restrict__ p, Grid g, Slot s,
                                                      const int8_t *mt, int64_t M, StepF h, double dt,
                                                      const double *lamTp, const double *LS, double *lamT,
                                                      NodeAdj L) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < M; base += stride) {
    const int64_t i = base + threadIdx.x;
    double v[5][4];
#pragma unroll
    for (int q = 0; q < 5; q++)
#pragma unroll
      for (int c = 0; c < 4; c++) v[q][c] = 0;
    int key = -1;
    if (i < M) {
      const float fx = s.x[i], fy = s.y[i];
      const double lp = lamTp[i];
      if (inside_strict(g, fx, fy) && mt[i] < 50) {
        const double x = fx, y = fy;
        const int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
        W4 w;
        nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
        key = m10 * g.Y + m20;
        const double e3 = isum(w, g.Y, h.tko);
        const float fk = s.T[i];
        double ld = 0, f = 0, lD = 0, D = 0, mkt = 1, mro = 1, mcp = 1;
        if (p->heatdif) {
          ld = lp + isum(w, g.Y, LS);
          const double celdx = g.gx[m10 + 1] - g.gx[m10], celdy = g.gy[m20 + 1] - g.gy[m20];
          mkt = isum(w, g.Y, h.kt);
          mcp = isum(w, g.Y, h.cp);
          mro = isum(w, g.Y, h.ro);
          D = p->heatdif * mkt / mro / mcp * 2.0 * (1.0 / celdx / celdx + 1.0 / celdy / celdy) * dt;
          double dfdD = 0;
          if (-D < -150.0) {
            f = 1.0 - exp(-150.0);
          } else {
            f = 1.0 - exp(-D);
            dfdD = exp(-D);
          }
          if (fk > 0) lD = ld * (e3 - (double)fk) * dfdD;
        }
        double le3, lt;
        if (fk > 0) {
          lt = lp - ld * f;
          le3 = -lp + ld * f;
        } else {  // markk <= 0 is replaced by e3 first (heatmart.c:541)
          lt = 0;
          le3 = 0;
        }
        lamT[i] = lt;
        const double lkt = lD * D / mkt, lro = -lD * D / mro, lcp = -lD * D / mcp;
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const double wc = corner_w(w, c);
          v[0][c] = wc * lp;  // e2: node T after the solve
          v[1][c] = wc * le3;
          v[2][c] = wc * lkt;
          v[3][c] = wc * lro;
          v[4][c] = wc * lcp;
        }
      } else {
        lamT[i] = lp;
      }
    }
    seg_scatter<5>(key, v, Dst<5>{{L.tnew, L.tk, L.kt, L.ro, L.cp}}, g.Y, lane);
  }
}

// ---------------------------------------------------------------- transposed heat rows
// y = A^T x for the rows of k_heat_apply (heatmart.c:406-432): equation rows
//   (A T)_m = T_m/dt - [(qx_m - qx_{m-Y})/xkf + (qy_m - qy_{m-1})/ykf] / (cp ro)_m
// and BC rows T_m - sum k T_nbr; the BC rows' neighbour columns are added by
// k_heat_apply_t_bc.
__global__ void k_heat_apply_t(Grid g, NodeF F, BcTab bc, double dt, const double *__restrict__ x,
                               double *__restrict__ y) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y, N = X * Y;
  if (j >= Y) return;
  const int n = i * Y + j;
  const double *gx = g.gx, *gy = g.gy, *kt = F.kt;
  auto eq = [&](int ii, int jj) { return ii >= 1 && ii <= X - 2 && jj >= 1 && jj <= Y - 2 && !bc.bondm[3 * N + ii * Y + jj]; };
  auto a = [&](int m) { return x[m] / F.cp[m] / F.ro[m]; };
  double r = 0;
  if (bc.bondm[3 * N + n]) {
    r = x[n];
  } else if (eq(i, j)) {
    const double xkf = (gx[i + 1] - gx[i - 1]) / 2.0, ykf = (gy[j + 1] - gy[j - 1]) / 2.0;
    const double kR = (kt[n] + kt[n + Y]) / 2.0 / (gx[i + 1] - gx[i]), kL = (kt[n - Y] + kt[n]) / 2.0 / (gx[i] - gx[i - 1]);
    const double kB = (kt[n] + kt[n + 1]) / 2.0 / (gy[j + 1] - gy[j]), kT = (kt[n - 1] + kt[n]) / 2.0 / (gy[j] - gy[j - 1]);
    r = x[n] / dt + a(n) * ((kR + kL) / xkf + (kB + kT) / ykf);
  }
  if (i >= 1 && eq(i - 1, j)) {  // n is the right neighbour of m = n - Y
    const int m = n - Y;
    const double xkf = (gx[i] - gx[i - 2]) / 2.0;
    r -= a(m) * ((kt[m] + kt[n]) / 2.0 / (gx[i] - gx[i - 1])) / xkf;
  }
  if (i + 1 <= X - 1 && eq(i + 1, j)) {
    const int m = n + Y;
    const double xkf = (gx[i + 2] - gx[i]) / 2.0;
    r -= a(m) * ((kt[n] + kt[m]) / 2.0 / (gx[i + 1] - gx[i])) / xkf;
  }
  if (j >= 1 && eq(i, j - 1)) {
    const int m = n - 1;
    const double ykf = (gy[j] - gy[j - 2]) / 2.0;
    r -= a(m) * ((kt[m] + kt[n]) / 2.0 / (gy[j] - gy[j - 1])) / ykf;
  }
  if (j + 1 <= Y - 1 && eq(i, j + 1)) {
    const int m = n + 1;
    const double ykf = (gy[j + 2] - gy[j]) / 2.0;
    r -= a(m) * ((kt[n] + kt[m]) / 2.0 / (gy[j + 1] - gy[j])) / ykf;
  }
  y[n] = r;
}

__global__ void k_heat_apply_t_bc(BcTab bc, int N, const double *x, double *y) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  int b = bc.bondm[3 * N + m];
  if (!b || x[m] == 0) return;
  for (int n1 = 0; n1 < 3; n1++) {
    int nb = bc.bondn[3 * b + n1];
    if (nb) aadd(&y[nb - 1 - 3 * N], -bc.bondv[4 * b + n1 + 1] * x[m]);
  }
}

// heat solve transposed, parameter part: with mu = A^-T lambda,
//   dJ/dp += mu^T (db/dp - dA/dp T),  b_m = sol1_m + T_old,m / dt,
//   sol1_m = (ht_m + shear heating_m) / (cp ro)_m   (heatmart.c:74-87, 406-432)
__global__ void k_adj_heat_nodes(Grid g, BcTab bc, StepF h, double dt, const double *__restrict__ mu, NodeAdj L) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y, N = X * Y;
  if (j >= Y) return;
  const int n = i * Y + j;
  const double *gx = g.gx, *gy = g.gy, *kt = h.kt, *T = h.tkn;
  auto eq = [&](int ii, int jj) { return ii >= 1 && ii <= X - 2 && jj >= 1 && jj <= Y - 2 && !bc.bondm[3 * N + ii * Y + jj]; };
  auto a = [&](int ii, int jj) {
    const int m = ii * Y + jj;
    return eq(ii, jj) ? mu[m] / h.cp[m] / h.ro[m] : 0.0;
  };
  auto xkf = [&](int ii) { return (gx[ii + 1] - gx[ii - 1]) / 2.0; };
  auto ykf = [&](int jj) { return (gy[jj + 1] - gy[jj - 1]) / 2.0; };
  if (eq(i, j)) {
    const double qR = (kt[n] + kt[n + Y]) / 2.0 * (T[n + Y] - T[n]) / (gx[i + 1] - gx[i]);
    const double qL = (kt[n - Y] + kt[n]) / 2.0 * (T[n] - T[n - Y]) / (gx[i] - gx[i - 1]);
    const double qB = (kt[n] + kt[n + 1]) / 2.0 * (T[n + 1] - T[n]) / (gy[j + 1] - gy[j]);
    const double qT = (kt[n - 1] + kt[n]) / 2.0 * (T[n] - T[n - 1]) / (gy[j] - gy[j - 1]);
    const double Lm = (qR - qL) / xkf(i) + (qB - qT) / ykf(j);
    const double am = mu[n] / h.cp[n] / h.ro[n];
    L.tk[n] += mu[n] / dt;
    L.ht[n] += am;
    L.cp[n] += -mu[n] * h.sol1[n] / h.cp[n] - am * Lm / h.cp[n];
    L.ro[n] += -mu[n] * h.sol1[n] / h.ro[n] - am * Lm / h.ro[n];
  }
  // conductivity: node n enters the arithmetic mean of its four edges; edge
  // e = (L, R) appears as +q_e/xkf_L in row L and -q_e/xkf_R in row R
  double lk = 0;
  if (i + 1 <= X - 1) {
    const double G = (T[n + Y] - T[n]) / (gx[i + 1] - gx[i]);
    const double ce = (i >= 1 ? a(i, j) / xkf(i) : 0.0) - (i + 1 <= X - 2 ? a(i + 1, j) / xkf(i + 1) : 0.0);
    lk += ce * G / 2.0;
  }
  if (i >= 1) {
    const double G = (T[n] - T[n - Y]) / (gx[i] - gx[i - 1]);
    const double ce = (i - 1 >= 1 ? a(i - 1, j) / xkf(i - 1) : 0.0) - (i <= X - 2 ? a(i, j) / xkf(i) : 0.0);
    lk += ce * G / 2.0;
  }
  if (j + 1 <= Y - 1) {
    const double G = (T[n + 1] - T[n]) / (gy[j + 1] - gy[j]);
    const double ce = (j >= 1 ? a(i, j) / ykf(j) : 0.0) - (j + 1 <= Y - 2 ? a(i, j + 1) / ykf(j + 1) : 0.0);
    lk += ce * G / 2.0;
  }
  if (j >= 1) {
    const double G = (T[n] - T[n - 1]) / (gy[j] - gy[j - 1]);
    const double ce = (j - 1 >= 1 ? a(i, j - 1) / ykf(j - 1) : 0.0) - (j <= Y - 2 ? a(i, j) / ykf(j) : 0.0);
    lk += ce * G / 2.0;
  }
  L.kt[n] += lk;
}

// transposed preconditioner ends: f = r; z = u (rho cp on equation rows) / sc
template <class T>
__global__ void k_heatT_pre(int N, const double *r, T *f) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < N) f[m] = T(r[m]);
}
template <class T>
__global__ void k_heatT_post(int N, const T *u, const double *sc, const double *ro, const double *cp,
                             const int8_t *mask, double *z) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < N) z[m] = (double)u[m] * (mask[m] ? 1.0 : ro[m] * cp[m]) / sc[m];
}
__global__ void k_vmul(int n, const double *a, const double *b, double *c) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) c[m] = a[m] * b[m];
}
__global__ void k_vadd(int n, const double *a, double *b) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) b[m] += a[m];
}


// ---------------------------------------------------------------- coupled adjoint: the flow chain
// viscalc's Arrhenius branch (markmart.c:1959-1977) at temperature mtk: the
// viscosity and its derivatives in T and in s (a factor on every markdh, at
// s = 1); both 0 where the exponent cap or a viscosity limit holds
__device__ __forceinline__ double visc_td(const i2e_params *p, int t, double mtk, double &dT, double &ds) {
  const int mm2 = t >= 100 ? t - 100 : t;
  const int ty = t < 20 ? mm2 : mm2 - 20;
  const double raw = viscalc_d(p, mtk, 0.0, ty, 0.0);
  dT = ds = 0;
  const i2e_rock &r = p->rock[ty];
  if (ty <= 20 && r.nu0 != 0 && r.dh > 0 && (r.ss == 0 || r.mm == 1.0)) {
    const double e1 = r.dh / (8.314 * mtk), unc = r.nu0 * exp(e1);
    const bool lim = e1 > 150.0 || unc <= r.n0 || unc >= r.n1 || unc <= p->nubeg || unc >= p->nuend;
    if (!lim) {
      ds = raw * e1;
      dT = -raw * e1 / mtk;
    }
  }
  return raw;
}

// the viscosity ronurecalc scattered (markmart.c:320-340): the stored markv
// when positive, else viscalc of the marker temperature; eta_outer beyond
// radmax (outer = true)
__device__ __forceinline__ double ronu_eta(const i2e_params *p, const Grid &g, float fx, float fy, float v, int t,
                                           double mtk, bool &outer) {
  double eta;
  if (v > 0) {
    eta = (double)v;
  } else {
    double dT, ds;
    eta = visc_td(p, t, mtk, dT, ds);
  }
  const double cx = (double)fx - g.xsize / 2.0, cy = (double)fy - g.ysize / 2.0;
  outer = sqrt(cx * cx + cy * cy) > g.xsize / 2.0 * p->radmax_frac;
  if (outer) eta = p->eta_outer;
  return eta;
}

// tent sums {sum w eta, sum w} of nu (basic nodes) and nd (cell centres) at
// the recorded positions and viscosities
__global__ void k_cvisc_sums(const i2e_params *__restrict__ p, Grid g, Slot s, const int8_t *mt, const float *eta_r,
                             int64_t M, double *S) {
  const int N = g.X * g.Y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const float fx = s.x[i], fy = s.y[i];
    const int t = mt[i];
    if (!(inside_strict(g, fx, fy) && s.T[i] > 0 && t < 50)) continue;
    bool outer;
    const double eta = ronu_eta(p, g, fx, fy, eta_r[i], t, (double)s.T[i], outer);
    const Tent T = tent(g, (double)fx, (double)fy);
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

// nu / nd normalisation and [nubeg, nuend] clamp transposed (k_ronu_nodes):
// l <- l / W where the node took the unclamped mean, 0 where clamped; a node
// without markers kept its previous value, so its adjoint passes to the
// previous step's field.  A value that is neither mean nor clamp was set by
// the contrast limit: flagged.
__global__ void k_cvisc_nodes(int N, double nubeg, double nuend, const double *nu, const double *nd,
                              const double *S, NodeAdj L, NodeAdj Lp, int has_prev, int *flag) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  for (int f = 0; f < 2; f++) {
    const double W = S[(2 * f + 1) * N + m];
    double *l = f ? L.nd : L.nu, *lp = f ? Lp.nd : Lp.nu;
    if (W == 0) {
      if (has_prev) lp[m] += l[m];
      l[m] = 0;
      continue;
    }
    const double v = S[2 * f * N + m] / W, vc = v < nubeg ? nubeg : (v > nuend ? nuend : v);
    const double have = f ? nd[m] : nu[m];
    if (dabs(have - vc) > 1e-9 * vc) atomicOr(flag, 1);
    l[m] = v == vc ? l[m] / W : 0.0;
  }
}

// the marker values ronurecalc scatters bilinearly (markmart.c:475-785 with
// densimod 0/1 and temperature/pressure-independent properties, which the
// adjoint requires): density, heat capacity, conductivity, radiogenic heat
struct MkVals {
  double ro, cp, kt, ht;
};
__device__ __forceinline__ MkVals mk_vals(const i2e_params *p, int t, float por_in, HeatK hk) {
  const int mm2 = t >= 100 ? t - 100 : t;
  const i2e_rock &r = p->rock[mm2];
  float por = por_in;
  if (mm2 == 0 || mm2 >= 7) por = 0.000f;
  if (por < 0.000f) por = 0.000f;
  if (por > 1.000f) por = 1.000f;
  MkVals v;
  v.ro = p->densimod == 0 ? r.ro : r.ro * (1.000 - (double)por);
  v.cp = r.cp;
  double mkt = r.kt;
  if (p->porous[mm2]) {
    const double kt0 = r.kt, pr0 = (double)por;
    if (por < 0.200f) mkt = kt0 * exp(-pr0 / 0.080);
    if (por >= 0.200f && por <= 0.400f)
      mkt = pow((pow((kt0 * exp(-pr0 / 0.080)), 4.000) + pow((kt0 * exp(-1.200 - (pr0 / 0.167))), 4.000)), 0.250);
    if (por > 0.400f) mkt = kt0 * exp(-1.200 - (pr0 / 0.167));
  }
  if (mkt <= 0.0010) mkt = 0.0010;
  if (mkt >= p->k_cutoff) mkt = p->k_cutoff;
  v.kt = mkt;
  v.ht = p->is_silicate[mm2] ? hk.sil * v.ro : (p->is_iron[mm2] ? hk.fe * v.ro : 0.0);
  return v;
}

// the node fields a ronurecalc produced (basic-node means after the
// normalisation; node T after tkrecalc)
struct RonuF {
  const double *tk, *ro, *cp, *kt, *ht;
};

// ronurecalc r transposed in the marker positions: with F_m = sum w_im f_i /
// W_m, dF_m/dx_i = (dw_im/dx)(f_i - F_m)/W_m, summed over the bilinear
// fields (T, rho, cp, k, H; L already divided by W) and the tent fields (nu,
// nd; L already divided by their weight sums); also the viscosity adjoint
// leta_i = sum l w over the marker's two tent nodes
__global__ void k_cadj_ronu_pos(const i2e_params *__restrict__ p, Grid g, Slot s, const int8_t *mt,
                                const float *por_a, const float *eta_r, int64_t M, RonuF F, const double *S,
                                NodeAdj L, HeatK hk, double *lX, double *leta) {
  const int Y = g.Y, N = g.X * g.Y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const float fx = s.x[i], fy = s.y[i];
    const int t = mt[i];
    if (!(inside_strict(g, fx, fy) && s.T[i] > 0 && t < 50)) {
      leta[i] = 0;
      continue;
    }
    const double x = fx, y = fy, mtk = (double)s.T[i];
    const int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
    W4 w;
    nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
    const double hx = g.gx[m10 + 1] - g.gx[m10], hy = g.gy[m20 + 1] - g.gy[m20];
    const MkVals v = mk_vals(p, t, por_a[i], hk);
    double px = 0, py = 0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int m = (m10 + a) * Y + m20 + b;
        const double wxa = a ? w.wx1 : w.wx0, wyb = b ? w.wy1 : w.wy0;
        const double q = L.tk[m] * (mtk - F.tk[m]) + L.ro[m] * (v.ro - F.ro[m]) + L.cp[m] * (v.cp - F.cp[m]) +
                         L.kt[m] * (v.kt - F.kt[m]) + L.ht[m] * (v.ht - F.ht[m]);
        px += (a ? 1.0 : -1.0) / hx * wyb * q;
        py += wxa * (b ? 1.0 : -1.0) / hy * q;
      }
    bool outer;
    const double eta = ronu_eta(p, g, fx, fy, eta_r[i], t, mtk, outer);
    double le = 0;
    // nearest basic node: w = 4 (ddx - 1/2)(ddy - 1/2)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const double ddx = a ? w.wx1 : w.wx0, ddy = b ? w.wy1 : w.wy0;
        if (ddx > 0.5 && ddy > 0.5) {
          const int n = (m10 + a) * Y + m20 + b;
          const double W = S[N + n];
          if (W > 0) {
            const double q = L.nu[n] * (eta - S[n] / W);
            px += 4.0 * (a ? 1.0 : -1.0) / hx * (ddy - 0.5) * q;
            py += 4.0 * (ddx - 0.5) * (b ? 1.0 : -1.0) / hy * q;
            le += L.nu[n] * 4.0 * (ddx - 0.5) * (ddy - 0.5);
          }
        }
      }
    // cell centre: ddx = 1 - |wx0 - 1/2|, d ddx / dx = sign(wx0 - 1/2) / hx
    {
      const double ddx = 1.0 - dabs(w.wx0 - 0.5), ddy = 1.0 - dabs(w.wy0 - 0.5);
      if (ddx > 0.5 && ddy > 0.5) {
        const int c = (m10 + 1) * Y + m20 + 1;
        const double W = S[3 * N + c];
        if (W > 0) {
          const double q = L.nd[c] * (eta - S[2 * N + c] / W);
          const double sx = w.wx0 > 0.5 ? 1.0 : (w.wx0 < 0.5 ? -1.0 : 0.0);
          const double sy = w.wy0 > 0.5 ? 1.0 : (w.wy0 < 0.5 ? -1.0 : 0.0);
          px += 4.0 * sx / hx * (ddy - 0.5) * q;
          py += 4.0 * (ddx - 0.5) * sy / hy * q;
          le += L.nd[c] * 4.0 * (ddx - 0.5) * (ddy - 0.5);
        }
      }
    }
    lX[2 * i] += px;
    lX[2 * i + 1] += py;
    leta[i] = outer ? 0.0 : le;
  }
}

// the initial viscosities that ronurecalc 0 formed by viscalc (markv <= 0)
__global__ void k_cadj_eta0(const i2e_params *__restrict__ p, Grid g, Slot s, const int8_t *mt, const float *eta_r,
                            int64_t M, const double *leta, double *ls) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    if (leta[i] == 0 || eta_r[i] > 0) continue;
    double dT, ds;
    visc_td(p, mt[i], (double)s.T[i], dT, ds);
    ls[i] += leta[i] * ds;
  }
}

// movemark n transposed (markmart.c:27-115): the viscosity it stored for
// ronurecalc n+1 (viscalc of the marker temperature after titerate) into
// the temperature and s adjoints, and the advection x' = x + dt v(x)
// reversed: lX (of x') becomes the adjoint of x, the node velocities'
// adjoints are scattered into lv (Stokes layout, 3 m + 1, 3 m + 2)
__global__ void k_cadj_move(const i2e_params *__restrict__ p, Grid g, Slot s, const int8_t *mt, const float *eta_n1,
                            const float *T_n1, const double *vx, const double *vy, double dt, int64_t M,
                            const double *leta, double *lX, double *lamTp, double *ls, double *lv) {
  const int Y = g.Y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const float fx = s.x[i], fy = s.y[i];
    const int t = mt[i];
    const int mm2 = t >= 100 ? t - 100 : t;
    if (!((fx >= 0 && fy >= 0 && (double)fx <= g.xsize && (double)fy <= g.ysize) || p->outgrid != 1)) continue;
    if ((unsigned)mm2 < (unsigned)I2E_MAXTYPES && p->immobile[mm2]) continue;
    const double le = leta[i];
    if (le != 0 && t < 50 && eta_n1[i] > 0 && T_n1[i] > 0) {
      double dT, ds;
      visc_td(p, t, (double)T_n1[i], dT, ds);
      ls[i] += le * ds;
      lamTp[i] += le * dT;
    }
    double ax = lX[2 * i], ay = lX[2 * i + 1];
    if (ax == 0 && ay == 0) continue;
    Rk4 R;
    rk4(p, g, vx, vy, (double)fx, (double)fy, dt, R);
    const double gx = dt * ax, gy = dt * ay;
    if (p->markmod == 1) {
      scatter_w(R.s[0].wx, Y, gx, lv, 1);
      scatter_w(R.s[0].wy, Y, gy, lv, 2);
      double px = 0, py = 0;
      interp_grad_t(g, R.s[0].wx, 0, +1, vx, gx, R.s[0].cx, R.s[0].cy, px, py);
      interp_grad_t(g, R.s[0].wy, +1, 0, vy, gy, R.s[0].cx, R.s[0].cy, px, py);
      lX[2 * i] = ax + px;
      lX[2 * i + 1] = ay + py;
      continue;
    }
    // stage k's position is x + v_{k-1} h_k: its adjoint reaches x and v_{k-1}
    double bx[4] = {gx / 6.0, gx / 3.0, gx / 3.0, gx / 6.0}, by[4] = {gy / 6.0, gy / 3.0, gy / 3.0, gy / 6.0};
    const double h[4] = {0.0, dt / 2.0, dt / 2.0, dt};
    for (int k = 3; k >= 0; k--) {
      const Interp &I = R.s[k];
      scatter_w(I.wx, Y, bx[k], lv, 1);
      scatter_w(I.wy, Y, by[k], lv, 2);
      double px = 0, py = 0;
      interp_grad_t(g, I.wx, 0, +1, vx, bx[k], I.cx, I.cy, px, py);
      interp_grad_t(g, I.wy, +1, 0, vy, by[k], I.cx, I.cy, px, py);
      ax += px;
      ay += py;
      if (k > 0) {
        bx[k - 1] += px * h[k];
        by[k - 1] += py * h[k];
      }
    }
    lX[2 * i] = ax;
    lX[2 * i + 1] = ay;
  }
}

// titerate n's antidiffusion sums S1 = sum w dtk1 over its markers
// (k_heat_markers_a), recomputed from the recorded temperatures and fields
__global__ __launch_bounds__(256) void k_cadj_heat_s1(const i2e_params *__restrict__ p, Grid g, Slot s,
                                                      const int8_t *mt, int64_t M, StepF h, double dt, double *S1) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < M; base += stride) {
    const int64_t i = base + threadIdx.x;
    double v[1][4] = {{0, 0, 0, 0}};
    int key = -1;
    if (i < M) {
      const float fx = s.x[i], fy = s.y[i];
      if (inside_strict(g, fx, fy) && mt[i] < 50) {
        const double x = fx, y = fy;
        const int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
        W4 w;
        nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
        key = m10 * g.Y + m20;
        const double e3 = isum(w, g.Y, h.tko);
        float fk = s.T[i];
        if (fk <= 0) fk = (float)e3;
        const double celdx = g.gx[m10 + 1] - g.gx[m10], celdy = g.gy[m20 + 1] - g.gy[m20];
        const double mkt = isum(w, g.Y, h.kt), mcp = isum(w, g.Y, h.cp), mro = isum(w, g.Y, h.ro);
        double d = -p->heatdif * mkt / mro / mcp * 2.0 * (1.0 / celdx / celdx + 1.0 / celdy / celdy) * dt;
        if (d < -150.0) d = -150.0;
        const double dtk1 = (e3 - (double)fk) * (1.0 - exp(d));
#pragma unroll
        for (int c = 0; c < 4; c++) v[0][c] = corner_w(w, c) * dtk1;
      }
    }
    seg_scatter<1>(key, v, Dst<1>{{S1}}, g.Y, lane);
  }
}

// gradient of a bilinear basic-node interpolant in (x, y)
__device__ __forceinline__ void bgrad(const W4 &w, int Y, double hx, double hy, const double *f, double &gx,
                                      double &gy) {
  const int m = w.i0 * Y + w.j0;
  const double f00 = f[m], f01 = f[m + 1], f10 = f[m + Y], f11 = f[m + Y + 1];
  gx = (w.wy0 * (f10 - f00) + w.wy1 * (f11 - f01)) / hx;
  gy = (w.wx0 * (f01 - f00) + w.wx1 * (f11 - f10)) / hy;
}

// titerate n's marker update transposed in the positions (heatmart.c:529-689):
//   T' = T + dtk1 + (e2 - e3) - sum_m w_m R_m,  dtk1 = (e3 - T)(1 - exp(-D)),
//   R_m = S1_m / S0_m (sums over the markers of w and w dtk1; k_cadj_heat_s1)
// through e2, e3, D's interpolated k, rho, cp and the weights of S0, S1 and
// of the correction: sum_m dw_m [LS_m (dtk1 - R_m) - lam' R_m]
__global__ void k_cadj_heat_pos(const i2e_params *__restrict__ p, Grid g, Slot s, const int8_t *mt, int64_t M,
                                StepF h, double dt, const double *lamTp, const double *LS, const double *R,
                                double *lX) {
  const int Y = g.Y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const float fx = s.x[i], fy = s.y[i];
    if (!(inside_strict(g, fx, fy) && mt[i] < 50)) continue;
    const double lp = lamTp[i];
    const double x = fx, y = fy;
    const int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
    W4 w;
    nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
    const double hx = g.gx[m10 + 1] - g.gx[m10], hy = g.gy[m20 + 1] - g.gy[m20];
    const double e3 = isum(w, Y, h.tko);
    const float fk0 = s.T[i];
    const double fk = fk0 > 0 ? (double)fk0 : e3;
    double px = 0, py = 0, gxv, gyv;
    bgrad(w, Y, hx, hy, h.tkn, gxv, gyv);  // e2
    px += lp * gxv;
    py += lp * gyv;
    double le3 = 0;
    if (p->heatdif) {
      const double ld = lp + isum(w, Y, LS);
      const double mkt = isum(w, Y, h.kt), mcp = isum(w, Y, h.cp), mro = isum(w, Y, h.ro);
      const double D = p->heatdif * mkt / mro / mcp * 2.0 * (1.0 / hx / hx + 1.0 / hy / hy) * dt;
      double f, dfdD = 0;
      if (-D < -150.0) {
        f = 1.0 - exp(-150.0);
      } else {
        f = 1.0 - exp(-D);
        dfdD = exp(-D);
      }
      const double dtk1 = (e3 - fk) * f;
      if (fk0 > 0) {
        le3 = -lp + ld * f;
        const double lD = ld * (e3 - fk) * dfdD;
        if (lD != 0) {
          double a, b;
          bgrad(w, Y, hx, hy, h.kt, a, b);
          px += lD * D * a / mkt;
          py += lD * D * b / mkt;
          bgrad(w, Y, hx, hy, h.ro, a, b);
          px -= lD * D * a / mro;
          py -= lD * D * b / mro;
          bgrad(w, Y, hx, hy, h.cp, a, b);
          px -= lD * D * a / mcp;
          py -= lD * D * b / mcp;
        }
      }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          const int m = (m10 + a) * Y + m20 + b;
          const double Rm = R[m];
          const double q = LS[m] * (dtk1 - Rm) - lp * Rm;
          px += (a ? 1.0 : -1.0) / hx * (b ? w.wy1 : w.wy0) * q;
          py += (a ? w.wx1 : w.wx0) * (b ? 1.0 : -1.0) / hy * q;
        }
    } else if (fk0 > 0) {
      le3 = -lp;
    }
    if (le3 != 0) {
      bgrad(w, Y, hx, hy, h.tko, gxv, gyv);
      px += le3 * gxv;
      py += le3 * gyv;
    }
    lX[2 * i] += px;
    lX[2 * i + 1] += py;
  }
}

// ---------------------------------------------------------------- coupled adjoint: gravity rows
// (movemart.c:661-749) transposed.  With b' = the interior rows' part of
// lgp, x = G^-1 b' is G_II^-1 lgp_I on the interior rows (G_II symmetric on
// uniform grids) and 0 on the pointwise rows; the pointwise rows then take
// lb = (lgp - G_PI^T x) / mu.
__device__ __forceinline__ bool grav_interior(int X, int Y, int i, int j) {
  return i >= 2 && i <= X - 2 && j >= 2 && j <= Y - 2;
}
__global__ void k_cgrav_rhs(Grid g, const double *lgp, double *b) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= g.Y) return;
  const int m = i * g.Y + j;
  b[m] = grav_interior(g.X, g.Y, i, j) ? lgp[m] : 0.0;
}
__global__ void k_cgrav_post(Grid g, GravGeo q, const double *lgp, const double *x, double *lb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  if (grav_interior(X, Y, i, j)) {
    lb[m] = x[m];
    return;
  }
  double v = lgp[m];
  if (i > 0 && j > 0) {
    const double *gx = g.gx, *gy = g.gy;
    // rows r of which this node is a neighbour, and its coefficient there
    if (i + 1 <= X - 1 && grav_interior(X, Y, i + 1, j)) {  // west neighbour of r = (i+1, j)
      const int ii = i + 1;
      const double dx02 = (gx[ii] - gx[ii - 2]) / 2.0, dx24 = (gx[ii + 1] - gx[ii - 1]) / 2.0;
      v -= 2.0 / dx02 / (dx02 + dx24) * x[m + Y];
    }
    if (i - 1 >= 0 && grav_interior(X, Y, i - 1, j)) {  // east neighbour of r = (i-1, j)
      const int ii = i - 1;
      const double dx02 = (gx[ii] - gx[ii - 2]) / 2.0, dx24 = (gx[ii + 1] - gx[ii - 1]) / 2.0;
      v -= 2.0 / dx24 / (dx02 + dx24) * x[m - Y];
    }
    if (j + 1 <= Y - 1 && grav_interior(X, Y, i, j + 1)) {
      const int jj = j + 1;
      const double dy12 = (gy[jj] - gy[jj - 2]) / 2.0, dy23 = (gy[jj + 1] - gy[jj - 1]) / 2.0;
      v -= 2.0 / dy12 / (dy12 + dy23) * x[m + 1];
    }
    if (j - 1 >= 0 && grav_interior(X, Y, i, j - 1)) {
      const int jj = j - 1;
      const double dy12 = (gy[jj] - gy[jj - 2]) / 2.0, dy23 = (gy[jj + 1] - gy[jj - 1]) / 2.0;
      v -= 2.0 / dy23 / (dy12 + dy23) * x[m - 1];
    }
  }
  lb[m] = v / q.mu;
}
// lb_m d b_m / d gxkoef (k_grav_rhs)
__global__ void k_cgrav_gk(Grid g, GravGeo q, const double *lb, double *t) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  double r = 0;
  if (i > 0 && j > 0) {
    const double d = grav_dist(g, i, j);
    if (i == 1 || i == X - 1 || j == 1 || j == Y - 1)
      r = -q.R * q.R / d * q.mu;
    else if (d > q.R)
      r = -q.R * q.R / d * q.mu * 1e+18;
  }
  t[m] = r * lb[m];
}
// the density adjoint: interior rows b_m = 4/1.5 pi G (mean of the cell's
// four nodes), and gxkoef = 4/3 pi R G rho_mean over the cells inside R
// (coef = lgxkoef 4/3 pi R G / area inside)
__global__ void k_cgrav_ro(Grid g, GravGeo q, const double *lb, double coef, double *lro) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y;
  if (j >= Y) return;
  const double gamma = 6.67384e-11, pival = 2.0 * asin(1.0);
  double acc = 0;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      const int ci = i + a, cj = j + b;
      if (ci < 1 || ci > X - 1 || cj < 1 || cj > Y - 1) continue;
      const int m = ci * Y + cj;
      if (grav_interior(X, Y, ci, cj)) acc += lb[m] * 4.0 / 1.5 * pival * gamma / 4.0;
      if (coef != 0 && grav_dist(g, ci, cj) < q.R)
        acc += coef * (g.gx[ci] - g.gx[ci - 1]) * (g.gy[cj] - g.gy[cj - 1]) / 4.0;
    }
  lro[i * Y + j] += acc;
}
}  // namespace i2e

using namespace i2e;

static inline dim3 nb2a(int X, int Y) { return dim3((Y + 255) / 256, X); }
static inline unsigned mgrid(int64_t M) { return nblk(M, 256) < 65535u ? nblk(M, 256) : 65535u; }

// Rheology and closures the reverse pass differentiates: temperature- and
// pressure-independent density and conductivity, no melting, no sintering.
static int adj_supported(const i2e_ctx *c) {
  const i2e_params &P = c->P;
  if (P.melt_on || P.porosity_on || P.adiabyn > 0 || P.soft_turb || P.magnet) {
    set_error("adjoint: melting, porosity evolution, adiabatic heating and the soft-turbulence conductivity "
              "are not differentiated");
    return I2E_ERR_UNSUPPORTED;
  }
  if (c->tbc_passes > 1) {
    set_error("adjoint: chained temperature BC rows are not differentiated");
    return I2E_ERR_UNSUPPORTED;
  }
  for (int t = 0; t < I2E_MAXTYPES; t++)
    if (P.immobile[t]) {
      set_error("adjoint: immobile rock types are not differentiated");
      return I2E_ERR_UNSUPPORTED;
    }
  for (int t = 0; t < I2E_MAXTYPES; t++) {
    const i2e_rock &r = P.rock[t];
    if (r.aa != 0 || r.bb != 0 || r.kf != 0 || r.kp != 0) {
      set_error("adjoint: rock type %d has a temperature/pressure-dependent density or conductivity", t);
      return I2E_ERR_UNSUPPORTED;
    }
  }
  return I2E_OK;
}

static int rec_ronu(i2e_ctx *c, AdjHist *h, int r) {
  const int64_t M = c->M;
  if (M) HIP_OK(hipMemcpyAsync(h->meta.p + (size_t)r * M, c->mk.v, sizeof(float) * M, hipMemcpyDeviceToDevice, c->s));
  HIP_OK(hipMemcpyAsync(h->tmax0.p, c->mk.tmax, sizeof(float) * M, hipMemcpyDeviceToDevice, c->s));
  ST_OK(i2e_ronurecalc(c));
  if (M) {
    hipLaunchKernelGGL(k_last_set, dim3(mgrid(M)), dim3(256), 0, c->s, M, h->tmax0.p, c->mk.tmax, r, h->last.p);
    HIP_OK(hipMemcpyAsync(h->mx.p + (size_t)r * M, c->mk.x, sizeof(float) * M, hipMemcpyDeviceToDevice, c->s));
    HIP_OK(hipMemcpyAsync(h->my.p + (size_t)r * M, c->mk.y, sizeof(float) * M, hipMemcpyDeviceToDevice, c->s));
    HIP_OK(hipMemcpyAsync(h->mT.p + (size_t)r * M, c->mk.k, sizeof(float) * M, hipMemcpyDeviceToDevice, c->s));
  }
  h->tsum[r] = c->timesum;
  return I2E_OK;
}

static Slot slot(const AdjHist *h, int r) {
  const size_t o = (size_t)r * h->M;
  return Slot{h->mx.p + o, h->my.p + o, h->mT.p + o};
}
static StepF stepf(const AdjHist *h, int n, int N) {
  const double *b = h->nodes.p + (size_t)n * 6 * N;
  return StepF{b, b + N, b + 2 * (size_t)N, b + 3 * (size_t)N, b + 4 * (size_t)N, b + 5 * (size_t)N};
}

extern "C" {

int i2e_record(i2e_ctx *c, int nsteps) {
  if (!c || nsteps < 1 || nsteps > 32000 || !c->setup_done) return I2E_ERR_ARG;
  ST_OK(adj_supported(c));
  const int N = c->N;
  const int64_t M = c->M;
  if (!c->adj) c->adj = new AdjHist();
  AdjHist *h = c->adj;
  h->valid = false;
  h->lx_valid = false;
  const size_t MS = (size_t)(nsteps + 1) * (size_t)M;
  ST_OK(h->mx.alloc(MS));
  ST_OK(h->my.alloc(MS));
  ST_OK(h->mT.alloc(MS));
  ST_OK(h->last.alloc(M ? M : 1));
  ST_OK(h->tmax0.alloc(M ? M : 1));
  ST_OK(h->nodes.alloc((size_t)nsteps * 6 * N));
  ST_OK(h->fnodes.alloc((size_t)nsteps * 7 * N));
  ST_OK(h->meta.alloc(MS ? MS : 1));
  h->sdt.assign(nsteps, 0.0);
  h->vlimited = 0;
  h->dt.assign(nsteps, 0.0);
  h->tsum.assign(nsteps + 1, 0.0);
  h->solved.assign(nsteps, 0);
  h->shear = false;
  h->limited = 0;
  h->M = M;
  if (M) HIP_OK(hipMemsetAsync(h->last.p, 0xff, sizeof(int16_t) * M, c->s));
  ST_OK(rec_ronu(c, h, 0));
  for (int n = 0; n < nsteps; n++) {
    // i2e_step's sequence, with the thermal inputs and outputs kept
    c->timestep = c->P.maxtmstep;
    ST_OK(i2e_maxvelstep(c, nullptr));
    int vlim = c->timestep != c->P.maxtmstep;
    ST_OK(i2e_viterate(c, 0));
    h->sdt[n] = c->stokes_dt;
    c->timestep = c->P.maxtmstep;
    ST_OK(i2e_maxvelstep(c, nullptr));
    if (vlim || c->timestep != c->P.maxtmstep) h->vlimited++;
    double *nb = h->nodes.p + (size_t)n * 6 * N;
    const size_t B = sizeof(double) * N;
    {
      double *fb = h->fnodes.p + (size_t)n * 7 * N;
      const double *src[7] = {c->F.nu, c->F.nd, c->F.gp, c->F.ht, c->F.vx, c->F.vy, c->F.pr};
      for (int k = 0; k < 7; k++)
        HIP_OK(hipMemcpyAsync(fb + (size_t)k * N, src[k], B, hipMemcpyDeviceToDevice, c->s));
    }
    HIP_OK(hipMemcpyAsync(nb, c->F.tk, B, hipMemcpyDeviceToDevice, c->s));
    HIP_OK(hipMemcpyAsync(nb + 2 * (size_t)N, c->F.ro, B, hipMemcpyDeviceToDevice, c->s));
    HIP_OK(hipMemcpyAsync(nb + 3 * (size_t)N, c->F.cp, B, hipMemcpyDeviceToDevice, c->s));
    HIP_OK(hipMemcpyAsync(nb + 4 * (size_t)N, c->F.kt, B, hipMemcpyDeviceToDevice, c->s));
    c->heat_solved = 0;
    if (c->P.frictyn > 0 && c->timesum > c->P.frictyn) h->shear = true;
    const double dt0 = c->timestep;
    if (c->timestep) ST_OK(i2e_titerate(c, 0));
    if (c->timestep != dt0) h->limited++;
    h->solved[n] = c->heat_solved;
    HIP_OK(hipMemcpyAsync(nb + (size_t)N, c->F.tk, B, hipMemcpyDeviceToDevice, c->s));
    if (c->heat_solved)
      HIP_OK(hipMemcpyAsync(nb + 5 * (size_t)N, c->sol1.p, B, hipMemcpyDeviceToDevice, c->s));
    else
      HIP_OK(hipMemsetAsync(nb + 5 * (size_t)N, 0, B, c->s));
    h->dt[n] = c->timestep;
    if (c->timestep) ST_OK(i2e_movemark(c));
    ST_OK(rec_ronu(c, h, n + 1));
    c->timesum += c->timestep;
  }
  HIP_OK(hipStreamSynchronize(c->s));
  h->K = nsteps;
  h->valid = true;
  return I2E_OK;
}

int i2e_replay(i2e_ctx *c, int nsteps) {
  if (!c || !c->setup_done) return I2E_ERR_ARG;
  AdjHist *h = c->adj;
  if (!h || !h->valid || nsteps < 1 || nsteps > h->K || c->M != h->M) {
    set_error("replay: no recorded trajectory of %d steps for %lld markers", nsteps, (long long)c->M);
    return I2E_ERR_ARG;
  }
  if (h->shear) {
    set_error("replay: the recorded run had shear heating (timesum > frictyn), which replay cannot reproduce");
    return I2E_ERR_UNSUPPORTED;
  }
  if (h->limited) {
    set_error("replay: titerate's heating-rate limit set the time step of %d recorded steps", h->limited);
    return I2E_ERR_UNSUPPORTED;
  }
  ST_OK(adj_supported(c));
  const int64_t M = c->M;
  const Grid g = c->grid();
  c->timesum = h->tsum[0];
  if (M) hipLaunchKernelGGL(k_replay_move, dim3(mgrid(M)), dim3(256), 0, c->s, g, c->mk, M, h->mx.p, h->my.p);
  ST_OK(i2e_ronurecalc(c));
  for (int n = 0; n < nsteps; n++) {
    c->timestep = h->dt[n];
    if (c->timestep) ST_OK(i2e_titerate(c, 0));
    if (c->timestep != h->dt[n]) {
      set_error("replay: titerate's heating-rate limit changed the time step of step %d", n);
      return I2E_ERR_UNSUPPORTED;
    }
    if (c->timestep && M) {
      const Slot s = slot(h, n + 1);
      hipLaunchKernelGGL(k_replay_move, dim3(mgrid(M)), dim3(256), 0, c->s, g, c->mk, M, s.x, s.y);
    }
    ST_OK(i2e_ronurecalc(c));
    c->timesum += c->timestep;
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->s));
  return I2E_OK;
}

static int peak_sums(i2e_ctx *c, double *smt, double *sm) {
  AdjHist *h = c->adj;
  if (!h) c->adj = h = new AdjHist();
  ST_OK(h->part.alloc(2 * RG + 2));
  hipLaunchKernelGGL(k_peak_part, dim3(RG), dim3(RB), 0, c->s, c->dP, c->grid(), c->mk, c->M, h->part.p);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RB), 0, c->s, h->part.p, h->part.p + 2 * RG);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(c->pin.p, h->part.p + 2 * RG, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  *smt = c->pin.p[0];
  *sm = c->pin.p[1];
  return I2E_OK;
}

int i2e_peak_t(i2e_ctx *c, double *J) {
  if (!c || !J) return I2E_ERR_ARG;
  double smt, sm;
  ST_OK(peak_sums(c, &smt, &sm));
  *J = sm > 0 ? smt / sm : 0.0;
  return I2E_OK;
}

int i2e_heat_apply_t_dev(i2e_ctx *c, const double *x, double *y) {
  if (!c || !x || !y) return I2E_ERR_ARG;
  hipLaunchKernelGGL(k_heat_apply_t, nb2a(c->X, c->Y), dim3(256), 0, c->s, c->grid(), c->F, c->bc, c->timestep, x, y);
  hipLaunchKernelGGL(k_heat_apply_t_bc, dim3(nblk(c->N, 256)), dim3(256), 0, c->s, c->bc, c->N, x, y);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->s));
  return I2E_OK;
}

}  // extern "C"

// transposed heat solve of step n: A^T mu = rhs (A the rows of that step)
static int heat_solve_t(i2e_ctx *c, const StepF &hf, double dt, const double *rhs, double *mu, double *nu,
                        double *tmp, int *its, double *rr) {
  const int X = c->X, Y = c->Y, N = c->N;
  const Grid g = c->grid();
  NodeF Fn = c->F;
  Fn.kt = const_cast<double *>(hf.kt);
  Fn.ro = const_cast<double *>(hf.ro);
  Fn.cp = const_cast<double *>(hf.cp);
  hipLaunchKernelGGL(k_heat_rowscale, nb2a(X, Y), dim3(256), 0, c->s, g, Fn, c->bc, dt, c->hsc.p);
  hipLaunchKernelGGL(k_heat_fields, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, hf.ro, hf.cp, dt, c->bondm.p,
                     c->hc.p, c->hmask.p, c->hdm.p);
  const bool mx = c->o.precision == I2E_PREC_MIXED;
  if (mx) {
    ST_OK(c->h32.refresh(hf.kt, c->hc.p, c->hmask.p, c->hdm.p, c->s));
    c->h32.nsmooth = c->o.nsmooth;
  } else {
    ST_OK(c->h64.refresh(hf.kt, c->hc.p, c->hmask.p, c->hdm.p, c->s));
    c->h64.nsmooth = c->o.nsmooth;
  }
  // (S A)^T nu = rhs with S = diag(hsc), then mu = S nu
  auto A = [&](const double *x, double *y) -> int {
    hipLaunchKernelGGL(k_vmul, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, c->hsc.p, x, tmp);
    hipLaunchKernelGGL(k_heat_apply_t, nb2a(X, Y), dim3(256), 0, c->s, g, Fn, c->bc, dt, tmp, y);
    hipLaunchKernelGGL(k_heat_apply_t_bc, dim3(nblk(N, 256)), dim3(256), 0, c->s, c->bc, N, tmp, y);
    return I2E_OK;
  };
  auto Mf = [&](const double *r, double *z) -> int {
    if (mx) {
      hipLaunchKernelGGL(k_heatT_pre<float>, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, r, c->h32.L[0].f.p);
      ST_OK(c->h32.vcycle(0, c->s));
      hipLaunchKernelGGL(k_heatT_post<float>, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, c->h32.L[0].u.p,
                         c->hsc.p, hf.ro, hf.cp, c->hmask.p, z);
    } else {
      hipLaunchKernelGGL(k_heatT_pre<double>, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, r, c->h64.L[0].f.p);
      ST_OK(c->h64.vcycle(0, c->s));
      hipLaunchKernelGGL(k_heatT_post<double>, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, c->h64.L[0].u.p,
                         c->hsc.p, hf.ro, hf.cp, c->hmask.p, z);
    }
    return I2E_OK;
  };
  HIP_OK(hipMemsetAsync(nu, 0, sizeof(double) * N, c->s));
  int st = c->fh.solve(A, Mf, rhs, nu, c->o.heat_rtol, c->o.maxit, c->s, its, rr);
  if (st) {
    set_error("adjoint heat solve: %d iterations, relative residual %.3e > %.1e", *its, *rr, c->o.heat_rtol);
    return st;
  }
  hipLaunchKernelGGL(k_vmul, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, c->hsc.p, nu, mu);
  return I2E_OK;
}

extern "C" int i2e_adjoint_peak_t(i2e_ctx *c, i2e_adjoint_out *out) {
  if (!c || !out) return I2E_ERR_ARG;
  AdjHist *h = c->adj;
  if (!h || !h->valid || c->M != h->M) {
    set_error("adjoint: no recorded trajectory (call i2e_record first)");
    return I2E_ERR_ARG;
  }
  ST_OK(adj_supported(c));
  if (h->limited) {
    set_error("adjoint: titerate's heating-rate limit (heatmart.c:366-387) set the time step of %d recorded "
              "steps; the gradient through that limit is not built (raise maxtkstep)", h->limited);
    return I2E_ERR_UNSUPPORTED;
  }
  const int N = c->N, X = c->X, Y = c->Y, K = h->K;
  const int64_t M = c->M;
  const Grid g = c->grid();
  const size_t MM = M ? (size_t)M : 1;
  ST_OK(h->lmax.alloc(MM));
  ST_OK(h->lamT.alloc(MM));
  ST_OK(h->lamTp.alloc(MM));
  ST_OK(h->lpor.alloc(MM));
  ST_OK(h->lh0.alloc(MM));
  ST_OK(h->L.alloc((size_t)2 * NADJ * N));
  ST_OK(h->nw.alloc((size_t)6 * N));
  memset(out, 0, sizeof *out);
  double smt, sm;
  ST_OK(peak_sums(c, &smt, &sm));
  if (!(sm > 0)) {
    set_error("adjoint: no silicate marker inside the grid");
    return I2E_ERR_ARG;
  }
  const double J = smt / sm;
  out->J = J;
  const unsigned gm = mgrid(M ? M : 1);
  HIP_OK(hipMemsetAsync(h->lamT.p, 0, sizeof(double) * MM, c->s));
  HIP_OK(hipMemsetAsync(h->lh0.p, 0, sizeof(double) * MM, c->s));
  HIP_OK(hipMemsetAsync(h->L.p, 0, sizeof(double) * 2 * NADJ * N, c->s));
  if (M) hipLaunchKernelGGL(k_adj_init, dim3(gm), dim3(256), 0, c->s, c->dP, g, c->mk, M, J, sm, h->lmax.p, h->lpor.p);
  double *W = h->nw.p, *S0 = W + N, *LS = W + 2 * (size_t)N, *mu = W + 3 * (size_t)N, *nu = W + 4 * (size_t)N,
         *tmp = W + 5 * (size_t)N;
  int cur = 0;
  double rrmax = 0;
  int itsum = 0;
  for (int r = K; r >= 0; r--) {
    NodeAdj Lc = node_adj(h->L.p + (size_t)cur * NADJ * N, N), Lp = node_adj(h->L.p + (size_t)(1 - cur) * NADJ * N, N);
    // ronurecalc r
    const Slot sr = slot(h, r);
    HIP_OK(hipMemsetAsync(W, 0, sizeof(double) * N, c->s));
    if (M) hipLaunchKernelGGL(k_adj_weights, dim3(gm), dim3(256), 0, c->s, g, sr, c->mk.t, M, W);
    hipLaunchKernelGGL(k_adj_tbc, dim3(nblk(N, 256)), dim3(256), 0, c->s, c->bc, N, Lc.tk);
    hipLaunchKernelGGL(k_adj_ronu_nodes, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, W, Lc, Lp, r > 0 ? 1 : 0);
    if (M)
      hipLaunchKernelGGL(k_adj_ronu_markers, dim3(gm), dim3(256), 0, c->s, c->dP, g, sr, c->mk.t, c->mk.por, M,
                         h->last.p, r, h->lmax.p, h->lamT.p, h->lamTp.p, h->lpor.p, h->lh0.p, Lc, h->tsum[r],
                         (r > 0 && h->dt[r - 1] != 0) ? 1 : 0);
    if (r == 0) break;
    const int n = r - 1;
    if (!h->solved[n]) {
      // titerate left node T and the markers as they were
      hipLaunchKernelGGL(k_vadd, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, Lp.tnew, Lp.tk);
      HIP_OK(hipMemcpyAsync(h->lamT.p, h->lamTp.p, sizeof(double) * MM, hipMemcpyDeviceToDevice, c->s));
    } else {
      const Slot sn = slot(h, n);
      const StepF hf = stepf(h, n, N);
      HIP_OK(hipMemsetAsync(S0, 0, sizeof(double) * 2 * N, c->s));
      if (M) {
        hipLaunchKernelGGL(k_adj_heat_mk1, dim3(gm), dim3(256), 0, c->s, g, sn, c->mk.t, M, h->lamTp.p, S0, LS);
        hipLaunchKernelGGL(k_adj_div, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, S0, LS);
        hipLaunchKernelGGL(k_adj_heat_mk2, dim3(gm), dim3(256), 0, c->s, c->dP, g, sn, c->mk.t, M, hf, h->dt[n],
                           h->lamTp.p, LS, h->lamT.p, Lp);
      }
      int its = 0;
      double rr = 0;
      ST_OK(heat_solve_t(c, hf, h->dt[n], Lp.tnew, mu, nu, tmp, &its, &rr));
      itsum += its;
      rrmax = rr > rrmax ? rr : rrmax;
      hipLaunchKernelGGL(k_adj_heat_nodes, nb2a(X, Y), dim3(256), 0, c->s, g, c->bc, hf, h->dt[n], mu, Lp);
    }
    // the next pass's ronurecalc reads what this step's titerate produced
    HIP_OK(hipMemsetAsync(Lc.tk, 0, sizeof(double) * NADJ * N, c->s));
    cur = 1 - cur;
  }
  HIP_OK(hipGetLastError());
  // deterministic sums of the per-marker gradients
  ST_OK(h->part.alloc(2 * RG + 2));
  hipLaunchKernelGGL(k_sum_part, dim3(RG), dim3(RB), 0, c->s, (int64_t)M, h->lh0.p, h->lpor.p, h->part.p);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RB), 0, c->s, h->part.p, h->part.p + 2 * RG);
  HIP_OK(hipMemcpyAsync(c->pin.p, h->part.p + 2 * RG, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  out->dJ_dh0 = c->pin.p[0];
  out->dJ_dphi0 = c->pin.p[1];
  out->steps = K;
  out->heat_iters = itsum;
  out->heat_relres_max = rrmax;
  return I2E_OK;
}

// ---------------------------------------------------------------- coupled adjoint: driver
// rheologies the flow part differentiates: Newtonian Arrhenius or constant,
// without activation volume, negative activation energy or power law
static int coupled_supported(i2e_ctx *c, const AdjHist *h) {
  const i2e_params &P = c->P;
  for (int t = 0; t < I2E_MAXTYPES; t++) {
    const i2e_rock &r = P.rock[t];
    if (r.nu0 != 0 && (r.dv != 0 || r.dh < 0 || (r.ss != 0 && r.mm != 1.0))) {
      set_error("coupled adjoint: rock type %d has an activation volume, a negative activation energy or a power "
                "law, which are not differentiated", t);
      return I2E_ERR_UNSUPPORTED;
    }
  }
  if (h->limited || h->vlimited) {
    set_error("coupled adjoint: the time step of %d recorded steps was set by titerate's heating-rate limit or "
              "by maxvelstep; the gradient through those limits is not built (raise maxtkstep and maxxystep)",
              h->limited + h->vlimited);
    return I2E_ERR_UNSUPPORTED;
  }
  if (h->shear) {
    set_error("coupled adjoint: the recorded run had shear heating (timesum > frictyn), whose flow dependence is "
              "not differentiated");
    return I2E_ERR_UNSUPPORTED;
  }
  // the transposed gravity solve uses the symmetry of the interior rows
  std::vector<double> gx(c->X), gy(c->Y);
  HIP_OK(hipMemcpy(gx.data(), c->gx.p, sizeof(double) * c->X, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(gy.data(), c->gy.p, sizeof(double) * c->Y, hipMemcpyDeviceToHost));
  auto uniform = [](const std::vector<double> &v) {
    const double d = (v.back() - v.front()) / (double)(v.size() - 1);
    for (size_t k = 0; k + 1 < v.size(); k++)
      if (std::fabs(v[k + 1] - v[k] - d) > 1e-9 * d) return false;
    return true;
  };
  if (!uniform(gx) || !uniform(gy)) {
    set_error("coupled adjoint: non-uniform grid lines are not supported (the transposed gravity solve needs "
              "symmetric interior rows)");
    return I2E_ERR_UNSUPPORTED;
  }
  return I2E_OK;
}

extern "C" int i2e_adjoint_peak_t_coupled(i2e_ctx *c, i2e_coupled_adjoint_out *out) {
  if (!c || !out) return I2E_ERR_ARG;
  AdjHist *h = c->adj;
  if (!h || !h->valid || c->M != h->M) {
    set_error("adjoint: no recorded trajectory (call i2e_record first)");
    return I2E_ERR_ARG;
  }
  ST_OK(adj_supported(c));
  ST_OK(coupled_supported(c, h));
  const int N = c->N, X = c->X, Y = c->Y, K = h->K;
  const int64_t M = c->M;
  const Grid g = c->grid();
  const size_t MM = M ? (size_t)M : 1, n3 = 3 * (size_t)N;
  ST_OK(h->lmax.alloc(MM));
  ST_OK(h->lamT.alloc(MM));
  ST_OK(h->lamTp.alloc(MM));
  ST_OK(h->lpor.alloc(MM));
  ST_OK(h->lh0.alloc(MM));
  ST_OK(h->lX.alloc(2 * MM));
  ST_OK(h->leta.alloc(MM));
  ST_OK(h->ls.alloc(MM));
  ST_OK(h->L.alloc((size_t)2 * NADJ * N));
  ST_OK(h->nw.alloc((size_t)6 * N));
  // flow work: S (4N), lv, mu, xpk (3N each), scratch (9N), lnu lnd lro lgp
  // (4N), gravity rhs / solution / adjoint / terms (4N), antidiffusion R (N)
  ST_OK(h->fw.alloc(4 * (size_t)N + 3 * n3 + 3 * n3 + 9 * (size_t)N));
  ST_OK(c->fflag.alloc(1));
  double *S = h->fw.p, *lv = S + 4 * (size_t)N, *mu3 = lv + n3, *xpk = mu3 + n3, *scr = xpk + n3,
         *lnu = scr + 3 * n3, *lnd = lnu + N, *lro = lnd + N, *lgp = lro + N, *gb = lgp + N, *gxs = gb + N,
         *glb = gxs + N, *gt = glb + N, *Rn = gt + N;
  memset(out, 0, sizeof *out);
  double smt, sm;
  ST_OK(peak_sums(c, &smt, &sm));
  if (!(sm > 0)) {
    set_error("adjoint: no silicate marker inside the grid");
    return I2E_ERR_ARG;
  }
  const double J = smt / sm;
  out->J = J;
  // gravity: the planet's area for the mean density (geometry only)
  const double xstpx = c->P.xsize / (double)(X - 1), ystpy = c->P.ysize / (double)(Y - 1);
  const GravGeo q{c->P.gykoef, c->gxkoef, 2.0 / xstpx / xstpx + 2 / ystpy / ystpy};
  double ronum = 0;
  {
    const int nbm = 256;
    hipLaunchKernelGGL(k_grav_mean, dim3(nbm), dim3(256), 0, c->s, g, h->nodes.p + 2 * (size_t)N, c->P.gykoef,
                       c->red.p);
    HIP_OK(hipMemcpyAsync(c->pin.p, c->red.p, 2 * nbm * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIP_OK(hipStreamSynchronize(c->s));
    for (int b = 0; b < nbm; b++) ronum += c->pin.p[2 * b + 1];
  }
  const unsigned gm = mgrid(M ? M : 1);
  HIP_OK(hipMemsetAsync(h->lamT.p, 0, sizeof(double) * MM, c->s));
  HIP_OK(hipMemsetAsync(h->lh0.p, 0, sizeof(double) * MM, c->s));
  HIP_OK(hipMemsetAsync(h->lX.p, 0, sizeof(double) * 2 * MM, c->s));
  HIP_OK(hipMemsetAsync(h->leta.p, 0, sizeof(double) * MM, c->s));
  HIP_OK(hipMemsetAsync(h->ls.p, 0, sizeof(double) * MM, c->s));
  HIP_OK(hipMemsetAsync(h->L.p, 0, sizeof(double) * 2 * NADJ * N, c->s));
  HIP_OK(hipMemsetAsync(c->fflag.p, 0, sizeof(int), c->s));
  if (M) hipLaunchKernelGGL(k_adj_init, dim3(gm), dim3(256), 0, c->s, c->dP, g, c->mk, M, J, sm, h->lmax.p, h->lpor.p);
  double *W = h->nw.p, *S0 = W + N, *LS = W + 2 * (size_t)N, *mu = W + 3 * (size_t)N, *nu = W + 4 * (size_t)N,
         *tmp = W + 5 * (size_t)N;
  int cur = 0;
  double rrh = 0, rrs = 0, rrg = 0;
  int64_t ith = 0, its_ = 0, itg = 0;
  const size_t B = sizeof(double) * N;
  for (int r = K; r >= 0; r--) {
    NodeAdj Lc = node_adj(h->L.p + (size_t)cur * NADJ * N, N), Lp = node_adj(h->L.p + (size_t)(1 - cur) * NADJ * N, N);
    const Slot sr = slot(h, r);
    const float *eta_r = h->meta.p + (size_t)r * h->M;
    // ronurecalc r
    HIP_OK(hipMemsetAsync(W, 0, B, c->s));
    HIP_OK(hipMemsetAsync(S, 0, 4 * B, c->s));
    if (M) {
      hipLaunchKernelGGL(k_adj_weights, dim3(gm), dim3(256), 0, c->s, g, sr, c->mk.t, M, W);
      hipLaunchKernelGGL(k_cvisc_sums, dim3(gm), dim3(256), 0, c->s, c->dP, g, sr, c->mk.t, eta_r, M, S);
    }
    hipLaunchKernelGGL(k_adj_tbc, dim3(nblk(N, 256)), dim3(256), 0, c->s, c->bc, N, Lc.tk);
    hipLaunchKernelGGL(k_adj_ronu_nodes, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, W, Lc, Lp, r > 0 ? 1 : 0);
    if (r < K) {
      const double *fr = h->fnodes.p + (size_t)r * 7 * N;
      hipLaunchKernelGGL(k_cvisc_nodes, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, c->P.nubeg, c->P.nuend, fr,
                         fr + N, (const double *)S, Lc, Lp, r > 0 ? 1 : 0, c->fflag.p);
    }
    if (M)
      hipLaunchKernelGGL(k_adj_ronu_markers, dim3(gm), dim3(256), 0, c->s, c->dP, g, sr, c->mk.t, c->mk.por, M,
                         h->last.p, r, h->lmax.p, h->lamT.p, h->lamTp.p, h->lpor.p, h->lh0.p, Lc, h->tsum[r],
                         (r > 0 && h->dt[r - 1] != 0) ? 1 : 0);
    if (r < K && M) {
      const StepF hr = stepf(h, r, N);
      const RonuF F{hr.tko, hr.ro, hr.cp, hr.kt, h->fnodes.p + ((size_t)r * 7 + 3) * N};
      hipLaunchKernelGGL(k_cadj_ronu_pos, dim3(gm), dim3(256), 0, c->s, c->dP, g, sr, c->mk.t, c->mk.por, eta_r, M,
                         F, (const double *)S, Lc, heat_per_kg(c->P, h->tsum[r]), h->lX.p, h->leta.p);
    } else {
      HIP_OK(hipMemsetAsync(h->leta.p, 0, sizeof(double) * MM, c->s));
    }
    if (r == 0) {
      if (M) hipLaunchKernelGGL(k_cadj_eta0, dim3(gm), dim3(256), 0, c->s, c->dP, g, sr, c->mk.t, eta_r, M,
                                (const double *)h->leta.p, h->ls.p);
      break;
    }
    const int n = r - 1;
    const Slot sn = slot(h, n);
    const double *fn = h->fnodes.p + (size_t)n * 7 * N;
    // movemark n
    HIP_OK(hipMemsetAsync(lv, 0, sizeof(double) * n3, c->s));
    if (M && h->dt[n] != 0)
      hipLaunchKernelGGL(k_cadj_move, dim3(gm), dim3(256), 0, c->s, c->dP, g, sn, c->mk.t, eta_r, sr.T, fn + 4 * (size_t)N,
                         fn + 5 * (size_t)N, h->dt[n], M, (const double *)h->leta.p, h->lX.p, h->lamTp.p, h->ls.p, lv);
    // titerate n
    const StepF hf = stepf(h, n, N);
    if (!h->solved[n]) {
      hipLaunchKernelGGL(k_vadd, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, Lp.tnew, Lp.tk);
      HIP_OK(hipMemcpyAsync(h->lamT.p, h->lamTp.p, sizeof(double) * MM, hipMemcpyDeviceToDevice, c->s));
    } else {
      HIP_OK(hipMemsetAsync(S0, 0, 2 * B, c->s));
      if (M) {
        hipLaunchKernelGGL(k_adj_heat_mk1, dim3(gm), dim3(256), 0, c->s, g, sn, c->mk.t, M, h->lamTp.p, S0, LS);
        hipLaunchKernelGGL(k_adj_div, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, S0, LS);
        if (c->P.heatdif) {
          HIP_OK(hipMemsetAsync(Rn, 0, B, c->s));
          hipLaunchKernelGGL(k_cadj_heat_s1, dim3(gm), dim3(256), 0, c->s, c->dP, g, sn, c->mk.t, M, hf, h->dt[n], Rn);
          hipLaunchKernelGGL(k_adj_div, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, S0, Rn);
        }
        hipLaunchKernelGGL(k_cadj_heat_pos, dim3(gm), dim3(256), 0, c->s, c->dP, g, sn, c->mk.t, M, hf, h->dt[n],
                           (const double *)h->lamTp.p, (const double *)LS, (const double *)Rn, h->lX.p);
        hipLaunchKernelGGL(k_adj_heat_mk2, dim3(gm), dim3(256), 0, c->s, c->dP, g, sn, c->mk.t, M, hf, h->dt[n],
                           h->lamTp.p, LS, h->lamT.p, Lp);
      }
      int its = 0;
      double rr = 0;
      ST_OK(heat_solve_t(c, hf, h->dt[n], Lp.tnew, mu, nu, tmp, &its, &rr));
      ith += its;
      rrh = rr > rrh ? rr : rrh;
      hipLaunchKernelGGL(k_adj_heat_nodes, nb2a(X, Y), dim3(256), 0, c->s, g, c->bc, hf, h->dt[n], mu, Lp);
    }
    // viterate n: the Stokes solve of that step's fields, transposed
    HIP_OK(hipMemcpyAsync(c->F.nu, fn, B, hipMemcpyDeviceToDevice, c->s));
    HIP_OK(hipMemcpyAsync(c->F.nd, fn + N, B, hipMemcpyDeviceToDevice, c->s));
    HIP_OK(hipMemcpyAsync(c->F.gp, fn + 2 * (size_t)N, B, hipMemcpyDeviceToDevice, c->s));
    HIP_OK(hipMemcpyAsync(c->F.ro, hf.ro, B, hipMemcpyDeviceToDevice, c->s));
    c->timestep = c->stokes_dt = h->sdt[n];
    c->have_prev = 0;
    ST_OK(stokes_prepare(c));
    {
      int its = 0;
      double rr = 0;
      ST_OK(stokes_solve_t(c, lv, mu3, scr, &its, &rr));
      its_ += its;
      rrs = rr > rrs ? rr : rrs;
    }
    hipLaunchKernelGGL(k_stokes_pack, nb2a(X, Y), dim3(256), 0, c->s, X, Y, fn + 6 * (size_t)N, fn + 4 * (size_t)N,
                       fn + 5 * (size_t)N, xpk);
    SOp<double> o = c->sop();
    o.dt = h->sdt[n];
    hipLaunchKernelGGL(k_stokes_field_adj, nb2a(X, Y), dim3(256), 0, c->s, o, c->bc, (const double *)xpk,
                       (const double *)mu3, lnu, lnd, lro, lgp);
    hipLaunchKernelGGL(k_vadd, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, lnu, Lp.nu);
    hipLaunchKernelGGL(k_vadd, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, lnd, Lp.nd);
    hipLaunchKernelGGL(k_vadd, dim3(nblk(N, 256)), dim3(256), 0, c->s, N, lro, Lp.ro);
    // gravity n: the potential's adjoint into the density's
    hipLaunchKernelGGL(k_cgrav_rhs, nb2a(X, Y), dim3(256), 0, c->s, g, (const double *)lgp, gb);
    {
      int its = 0;
      double rr = 0;
      ST_OK(gravity_solve_rhs(c, gb, gxs, &its, &rr));
      itg += its;
      rrg = rr > rrg ? rr : rrg;
    }
    hipLaunchKernelGGL(k_cgrav_post, nb2a(X, Y), dim3(256), 0, c->s, g, q, (const double *)lgp, (const double *)gxs,
                       glb);
    hipLaunchKernelGGL(k_cgrav_gk, nb2a(X, Y), dim3(256), 0, c->s, g, q, (const double *)glb, gt);
    ST_OK(h->part.alloc(2 * RG + 2));
    hipLaunchKernelGGL(k_sum_part, dim3(RG), dim3(RB), 0, c->s, (int64_t)N, gt, gt, h->part.p);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RB), 0, c->s, h->part.p, h->part.p + 2 * RG);
    HIP_OK(hipMemcpyAsync(c->pin.p, h->part.p + 2 * RG, sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIP_OK(hipStreamSynchronize(c->s));
    const double gamma = 6.67384e-11, pival = 2.0 * asin(1.0);
    const double coef = ronum > 0 ? c->pin.p[0] * 4.0 / 3.0 * pival * c->P.gykoef * gamma / ronum : 0.0;
    hipLaunchKernelGGL(k_cgrav_ro, nb2a(X, Y), dim3(256), 0, c->s, g, q, (const double *)glb, coef, Lp.ro);
    HIP_OK(hipMemsetAsync(Lc.tk, 0, sizeof(double) * NADJ * N, c->s));
    cur = 1 - cur;
  }
  HIP_OK(hipGetLastError());
  c->stokes_valid = 0;
  c->have_prev = 0;
  ST_OK(h->part.alloc(2 * RG + 2));
  hipLaunchKernelGGL(k_sum_part, dim3(RG), dim3(RB), 0, c->s, (int64_t)M, h->lh0.p, h->lpor.p, h->part.p);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RB), 0, c->s, h->part.p, h->part.p + 2 * RG);
  HIP_OK(hipMemcpyAsync(c->pin.p, h->part.p + 2 * RG, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  out->dJ_dh0 = c->pin.p[0];
  out->dJ_dphi0 = c->pin.p[1];
  hipLaunchKernelGGL(k_sum_part, dim3(RG), dim3(RB), 0, c->s, (int64_t)M, h->ls.p, h->ls.p, h->part.p);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RB), 0, c->s, h->part.p, h->part.p + 2 * RG);
  HIP_OK(hipMemcpyAsync(c->pin.p, h->part.p + 2 * RG, sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipMemcpyAsync(c->pin.p + 2, c->fflag.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  out->dJ_dEscale = c->pin.p[0];
  int fl;
  memcpy(&fl, c->pin.p + 2, sizeof fl);
  if (fl) {
    set_error("coupled adjoint: the viscosity contrast limit (nucontr) set some node viscosities; not differentiated");
    return I2E_ERR_UNSUPPORTED;
  }
  out->steps = K;
  out->heat_iters = ith;
  out->stokes_iters = its_;
  out->grav_iters = itg;
  out->heat_relres_max = rrh;
  out->stokes_relres_max = rrs;
  out->grav_relres_max = rrg;
  h->lx_valid = true;
  return I2E_OK;
}

// dJ/d(initial marker positions) of the last coupled adjoint, in the
// caller's marker order
extern "C" int i2e_adjoint_marker_positions(i2e_ctx *c, double *dJ_dx, double *dJ_dy) {
  if (!c || !dJ_dx || !dJ_dy) return I2E_ERR_ARG;
  AdjHist *h = c->adj;
  const int64_t M = c->M;
  if (!h || !h->valid || !h->lx_valid || h->M != M) {
    set_error("adjoint: no coupled adjoint of the recorded trajectory (call i2e_adjoint_peak_t_coupled first)");
    return I2E_ERR_ARG;
  }
  if (!M) return I2E_OK;
  std::vector<double> l(2 * (size_t)M);
  std::vector<int> orig;
  HIP_OK(hipMemcpy(l.data(), h->lX.p, sizeof(double) * 2 * M, hipMemcpyDeviceToHost));
  if (c->permuted) {
    orig.resize(M);
    HIP_OK(hipMemcpy(orig.data(), c->morig.p, sizeof(int) * M, hipMemcpyDeviceToHost));
  }
  for (int64_t i = 0; i < M; i++) {
    const int64_t k = c->permuted ? orig[i] : i;
    dJ_dx[k] = l[2 * i];
    dJ_dy[k] = l[2 * i + 1];
  }
  return I2E_OK;
}
