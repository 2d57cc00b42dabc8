// mg.hip -- geometric multigrid for the Stokes velocity block and for the
// scalar (heat, gravity) operators.  See mg.hpp for the design.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <type_traits>

#include "mg.hpp"

#include <map>
#include <hip/hip_cooperative_groups.h>
#include <mutex>
#include <tuple>

namespace i2e {

static inline dim3 blk2() { return dim3(64, 4); }
static inline dim3 grd2(int X, int Y) { return dim3((Y + 63) / 64, (X + 3) / 4); }

#define IDX2                                                   \
  const int j = blockIdx.x * blockDim.x + threadIdx.x;         \
  const int i = blockIdx.y * blockDim.y + threadIdx.y;

std::vector<int> coarse_lines(int n) {
  std::vector<int> v;
  for (int i = 0; i < n; i += 2) v.push_back(i);
  if (v.back() != n - 1) v.push_back(n - 1);
  return v;
}

int build_geometry(const double *gx, const double *gy, int X, int Y, int coarse_max, std::vector<GeoLevel> &geo) {
  geo.clear();
  GeoLevel g0;
  g0.X = X;
  g0.Y = Y;
  g0.hx.assign(gx, gx + X);
  g0.hy.assign(gy, gy + Y);
  geo.push_back(std::move(g0));
  while (true) {
    GeoLevel &f = geo.back();
    if (std::min(f.X, f.Y) - 1 <= coarse_max || std::min(f.X, f.Y) < 9) break;
    std::vector<int> ci = coarse_lines(f.X), cj = coarse_lines(f.Y);
    f.fi = ci;
    f.fj = cj;
    GeoLevel c;
    c.X = (int)ci.size();
    c.Y = (int)cj.size();
    for (int a : ci) c.hx.push_back(f.hx[a]);
    for (int a : cj) c.hy.push_back(f.hy[a]);
    geo.push_back(std::move(c));
  }
  for (auto &g : geo) {
    ST_OK(g.gx.alloc(g.X));
    ST_OK(g.gy.alloc(g.Y));
    HIP_OK(hipMemcpy(g.gx.p, g.hx.data(), g.X * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(g.gy.p, g.hy.data(), g.Y * sizeof(double), hipMemcpyHostToDevice));
  }
  return I2E_OK;
}

int Map1::build(const std::vector<double> &cf, const std::vector<double> &cc) {
  nf = (int)cf.size();
  nc = (int)cc.size();
  std::vector<int> hi0(nf), hrb(nc, nf), hre(nc, 0);
  std::vector<double> hw(nf);
  for (int k = 0; k < nf; k++) {
    int I = (int)(std::upper_bound(cc.begin(), cc.end(), cf[k]) - cc.begin()) - 1;
    if (I < 0) I = 0;
    if (I > nc - 2) I = nc - 2;
    double w = (cf[k] - cc[I]) / (cc[I + 1] - cc[I]);
    if (w < 0) w = 0;
    if (w > 1) w = 1;
    hi0[k] = I;
    hw[k] = w;
    if (w < 1) {
      hrb[I] = std::min(hrb[I], k);
      hre[I] = std::max(hre[I], k + 1);
    }
    if (w > 0) {
      hrb[I + 1] = std::min(hrb[I + 1], k);
      hre[I + 1] = std::max(hre[I + 1], k + 1);
    }
  }
  for (int I = 0; I < nc; I++)
    if (hre[I] <= hrb[I]) hrb[I] = hre[I] = 0;
  // normalised restriction weights (separable: the 2-D normalisation is the
  // product of the 1-D ones)
  std::vector<double> hrw((size_t)nc * RW, 0.0);
  for (int I = 0; I < nc; I++) {
    if (hre[I] - hrb[I] > RW) {
      set_error("multigrid restriction gathers %d fine lines (at most %d)", hre[I] - hrb[I], RW);
      return I2E_ERR_UNSUPPORTED;
    }
    double sum = 0;
    for (int k = hrb[I]; k < hre[I]; k++) {
      const double wk = (hi0[k] == I ? 1.0 - hw[k] : 0.0) + (hi0[k] + 1 == I ? hw[k] : 0.0);
      hrw[(size_t)I * RW + (k - hrb[I])] = wk;
      sum += wk;
    }
    for (int a = 0; a < RW; a++) hrw[(size_t)I * RW + a] = sum > 0 ? hrw[(size_t)I * RW + a] / sum : 0.0;
  }
  ST_OK(rw.alloc((size_t)nc * RW));
  HIP_OK(hipMemcpy(rw.p, hrw.data(), (size_t)nc * RW * sizeof(double), hipMemcpyHostToDevice));
  ST_OK(i0.alloc(nf));
  ST_OK(w.alloc(nf));
  ST_OK(rb.alloc(nc));
  ST_OK(re.alloc(nc));
  HIP_OK(hipMemcpy(i0.p, hi0.data(), nf * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(w.p, hw.data(), nf * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(rb.p, hrb.data(), nc * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(re.p, hre.data(), nc * sizeof(int), hipMemcpyHostToDevice));
  this->hrb = hrb;
  this->hre = hre;
  return I2E_OK;
}

int Map1::tile_span(int t) const {
  int worst = 0;
  for (int c0 = 0; c0 < nc; c0 += t) {
    int a0 = nf, a1 = 0;
    for (int c = c0; c < nc && c < c0 + t; c++)
      if (hre[c] > hrb[c]) {
        a0 = std::min(a0, hrb[c]);
        a1 = std::max(a1, hre[c]);
      }
    worst = std::max(worst, a1 - a0);
  }
  return worst;
}

__device__ __forceinline__ double wgt(const MapView &m, int k, int I) {
  int a = m.i0[k];
  double w = m.w[k];
  return (a == I ? 1.0 - w : 0.0) + (a + 1 == I ? w : 0.0);
}

// ---------------------------------------------------------------- generic kernels
template <class T>
__global__ void k_cast(int n, const double *a, T *b) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = T(a[t]);
}
template <class T>
__global__ void k_zero(int n, T *a) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = T(0);
}

// node field restriction through the node maps; geometric mean if lg
template <class T>
__global__ void k_restrict_node(int Xc, int Yc, int Yf, MapView mx, MapView my, const T *f, T *c, int lg) {
  IDX2;
  if (i >= Xc || j >= Yc) return;
  double s = 0, ws = 0;
  for (int k = mx.rb[i]; k < mx.re[i]; k++) {
    double wx = wgt(mx, k, i);
    if (wx == 0) continue;
    for (int l = my.rb[j]; l < my.re[j]; l++) {
      double w = wx * wgt(my, l, j);
      if (w == 0) continue;
      double v = (double)f[k * Yf + l];
      s += w * (lg ? log(v) : v);
      ws += w;
    }
  }
  double r = ws > 0 ? s / ws : 0.0;
  c[i * Yc + j] = T(lg ? exp(r) : r);
}

// cell field restriction: average over the fine cells inside coarse cell
// (cells are indexed by their lower-right node, i,j >= 1)
template <class T>
__global__ void k_restrict_cell(int Xc, int Yc, int Yf, const int *fi, const int *fj, const T *f, T *c, int lg) {
  IDX2;
  if (i >= Xc || j >= Yc) return;
  if (i == 0 || j == 0) {
    c[i * Yc + j] = f[fi[i] * Yf + fj[j]];
    return;
  }
  double s = 0, ws = 0;
  for (int a = fi[i - 1] + 1; a <= fi[i]; a++)
    for (int b = fj[j - 1] + 1; b <= fj[j]; b++) {
      double v = (double)f[a * Yf + b];
      s += lg ? log(v) : v;
      ws += 1;
    }
  double r = s / ws;
  c[i * Yc + j] = T(lg ? exp(r) : r);
}

// viscosity restriction: maximum over a window of +-rad fine lines around the
// fine point that carries the coarse node (cells: around their lower-right
// fine cell).  Local averages (geometric, arithmetic, harmonic) make coarse
// rows next to stiff regions much softer than the Galerkin operator R A P,
// and the coarse correction then diverges.
template <class T>
__global__ void k_restrict_max(int Xc, int Yc, int Xf, int Yf, const int *fi, const int *fj, int rad, const T *f, T *c) {
  IDX2;
  if (i >= Xc || j >= Yc) return;
  const int a0 = fi[i], b0 = fj[j];
  T v = f[a0 * Yf + b0];
  for (int a = a0 - rad; a <= a0 + rad; a++) {
    if (a < 0 || a >= Xf) continue;
    for (int b = b0 - rad; b <= b0 + rad; b++) {
      if (b < 0 || b >= Yf) continue;
      T u = f[a * Yf + b];
      v = u > v ? u : v;
    }
  }
  c[i * Yc + j] = v;
}

// ---------------------------------------------------------------- velocity rows
template <class T, class A>
__device__ __forceinline__ T vrow(const VLev<T> &L, const A &v, int comp, int i, int j) {
  const int Y = L.o.Y, m = i * Y + j;
  if (comp == 1) {
    int k = L.kx[m];
    if (k == 0) return xmom(L.o, v, i, j, false);
    T s = v.vx(i, j);
    if (k == 1 && L.nbx[m] >= 0) s -= L.kkx[m] * v.vx(L.nbx[m] / Y, L.nbx[m] % Y);
    return T(L.o.mu) * s;
  } else {
    int k = L.ky[m];
    if (k == 0) return ymom(L.o, v, i, j, false);
    T s = v.vy(i, j);
    if (k == 1 && L.nby[m] >= 0) s -= L.kky[m] * v.vy(L.nby[m] / Y, L.nby[m] % Y);
    return T(L.o.mu) * s;
  }
}

template <class T>
__global__ void k_vdiag(VLev<T> L, T *dx, T *dy) {
  IDX2;
  const int X = L.o.X, Y = L.o.Y;
  if (i >= X || j >= Y) return;
  const int m = i * Y + j;
  dx[m] = vrow(L, VUnit<T>{1, m, Y}, 1, i, j);
  dy[m] = vrow(L, VUnit<T>{2, m, Y}, 2, i, j);
}

template <class T>
__global__ void k_vresid(VLev<T> L, const T *ux, const T *uy, const T *fx, const T *fy, T *rx, T *ry) {
  IDX2;
  const int X = L.o.X, Y = L.o.Y;
  if (i >= X || j >= Y) return;
  const int m = i * Y + j;
  VSplit<T> v{ux, uy, nullptr, Y};
  rx[m] = L.kx[m] == 0 ? fx[m] - vrow(L, v, 1, i, j) : T(0);
  ry[m] = L.ky[m] == 0 ? fy[m] - vrow(L, v, 2, i, j) : T(0);
}

// coarse rhs = normalised transpose interpolation of the fine residual
// (fine BC rows were zeroed in k_vresid)
template <class T>
__device__ __forceinline__ T restrict_rw(const T *r, int Yf, const MapView &mx, const MapView &my, int i, int j) {
  const int k0 = mx.rb[i], l0 = my.rb[j];
  double wx[RW], wy[RW];
#pragma unroll
  for (int a = 0; a < RW; a++) {
    wx[a] = mx.rw[i * RW + a];
    wy[a] = my.rw[j * RW + a];
  }
  double s = 0;
#pragma unroll
  for (int a = 0; a < RW; a++)
#pragma unroll
    for (int b = 0; b < RW; b++)
      if (wx[a] != 0 && wy[b] != 0) s += wx[a] * wy[b] * (double)r[(k0 + a) * Yf + l0 + b];
  return T(s);
}

// coarse rhs = normalised transpose interpolation of the fine residual
// (fine pointwise rows carry zero residual)
template <class T>
__global__ void k_vrestrict(int Xc, int Yc, int Yf, MapView xn, MapView xm, MapView yn, MapView ym, const int8_t *ckx,
                            const int8_t *cky, const T *rx, const T *ry, T *fx, T *fy) {
  IDX2;
  if (i >= Xc || j >= Yc) return;
  const int m = i * Yc + j;
  // Vx: x at nodes, y at mids (only j <= Yc-2 carry Vx); Vy the other way round
  fx[m] = (ckx[m] == 0 && j <= Yc - 2) ? restrict_rw(rx, Yf, xn, ym, i, j) : T(0);
  fy[m] = (cky[m] == 0 && i <= Xc - 2) ? restrict_rw(ry, Yf, xm, yn, i, j) : T(0);
}

// dense matrix of the coarsest level, column c = A e_c, rows/cols ordered
// [Vx(0..N-1), Vy(0..N-1)]; stored into the left half of an n x 2n array
template <class T>
__global__ void k_vdense(VLev<T> L, int n, double *aug) {
  int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (c >= n) return;
  const int N = n / 2, Y = L.o.Y;
  int comp = r < N ? 1 : 2, m = r % N;
  VUnit<T> e{c < N ? 1 : 2, c % N, Y};
  aug[(size_t)r * 2 * n + c] = (double)vrow(L, e, comp, m / Y, m % Y);
  aug[(size_t)r * 2 * n + n + c] = (r == c) ? 1.0 : 0.0;
}

// Gauss-Jordan inversion with partial pivoting in one workgroup, on the
// n x 2n augmented matrix held in LDS (n <= GJ_LDS_N).  Three barriers per
// pivot: (1) column p is staged and the pivot found by a lane-shuffle
// maximum over the first two wavefronts (largest |a|, lowest row on ties);
// (2) rows p and q are exchanged with the new row p divided by the pivot;
// (3) every other row r subtracts (its staged column-p entry) x row p, which
// also leaves exact zeros in column p (f - f * 1).
constexpr int GJ_LDS_N = 90;
__global__ __launch_bounds__(1024) void k_gauss_jordan_lds(int n, const double *ag, double *inv) {
  __shared__ double a[GJ_LDS_N * 2 * GJ_LDS_N];
  __shared__ double colp[GJ_LDS_N], sv[2], sd[2];
  __shared__ int si[2];
  const int t = threadIdx.x, nt = blockDim.x, w = 2 * n;
  for (int idx = t; idx < n * w; idx += nt) a[idx] = ag[idx];
  // elimination layout: thread t updates column t % w of rows t / w + k * rpp
  const int rpp = nt / w, ce = t % w, re = t / w;
  __syncthreads();
  for (int p = 0; p < n; p++) {
    if (t < n) colp[t] = a[t * w + p];
    if (t < 128) {
      const int r = p + t;
      double best = -1, val = 0;
      int bi = n;
      if (r < n) {
        val = a[r * w + p];
        best = fabs(val);
        bi = r;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64), ov = __shfl_xor(val, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          val = ov;
          bi = oi;
        }
      }
      if ((t & 63) == 0) {
        sv[t >> 6] = best;
        sd[t >> 6] = val;
        si[t >> 6] = bi;
      }
    }
    __syncthreads();
    const bool w1 = sv[1] > sv[0];  // rows of the second wavefront lie below: ties stay in the first
    const int q = w1 ? si[1] : si[0];
    const double d = w1 ? sd[1] : sd[0];
    if (t < w) {
      const double x = a[p * w + t], y = a[q * w + t];
      a[p * w + t] = y / d;
      if (q != p) a[q * w + t] = x;
    }
    __syncthreads();
    if (re < rpp) {
      const double pr = a[p * w + ce];
      for (int r = re; r < n; r += rpp) {
        if (r == p) continue;
        const double f = r == q ? colp[p] : colp[r];
        a[r * w + ce] -= f * pr;
      }
    }
    __syncthreads();
  }
  for (int idx = t; idx < n * n; idx += nt) inv[idx] = a[(idx / n) * w + n + idx % n];
}

// In-place Gauss-Jordan inversion with partial pivoting for n too large for
// LDS: the n x n matrix lives in the registers of one 1024-thread workgroup
// (thread (g, c) holds rows g*RPG .. g*RPG+RPG-1 of column c).  Two barriers
// per pivot k: (A) the threads of column k stage it (double-buffered) and
// search their rows for the pivot, the threads of row k stage that row;
// (B) every thread picks the pivot q (largest |a|, lowest row on ties), the
// holders of row q form the new row k = row q / pivot into LDS and take the
// old row k in its place; (C) the holders of row k take the new row and all
// other rows subtract (their staged column-k entry) x row k, the column k
// entries becoming the inverse's (0 - f / pivot).  The column swaps implied
// by the row pivots are undone when the inverse is written out.
constexpr int GJR_T = 1024, GJR_R = 32, GJR_N = 192;
__global__ __launch_bounds__(GJR_T) void k_gauss_jordan_reg(int n, const double *ag, double *inv) {
  __shared__ double rowb[GJR_N], rowx[GJR_N], colb[2][GJR_N], pv[2 * GJR_N], sval[GJR_T / 32], sdv[GJR_T / 32];
  __shared__ int sidx[GJR_T / 32], piv[GJR_N], pinv[GJR_N];
  const int t = threadIdx.x, G = GJR_T / n, rpg = (n + G - 1) / G;
  const int g = t / n, c = t % n;
  const bool act = g < G;
  double a[GJR_R];
#pragma unroll
  for (int k2 = 0; k2 < GJR_R; k2++) {
    const int r = g * rpg + k2;
    a[k2] = (act && k2 < rpg && r < n) ? ag[(size_t)r * 2 * n + c] : 0.0;
  }
  for (int k = 0; k < n; k++) {
    const int gk = k / rpg, sk = k % rpg, bf = k & 1;
    // (A) stage column k and search it; stage row k
    if (act && c == k) {
      double best = -1, bv = 0;
      int bi = k;
#pragma unroll
      for (int k2 = 0; k2 < GJR_R; k2++) {
        const int r = g * rpg + k2;
        if (k2 < rpg && r < n) {
          colb[bf][r] = a[k2];
          if (r >= k && fabs(a[k2]) > best) {
            best = fabs(a[k2]);
            bv = a[k2];
            bi = r;
          }
        }
      }
      sval[g] = best;
      sdv[g] = bv;
      sidx[g] = bi;
    }
    if (act && g == gk) {
      double x = 0;
#pragma unroll
      for (int k2 = 0; k2 < GJR_R; k2++) x = k2 == sk ? a[k2] : x;
      rowx[c] = x;
    }
    __syncthreads();
    // (B) pivot row q: row k := row q / pivot (into rowb), row q := old row k
    double best = -1, d = 0;
    int q = k;
    for (int gg = 0; gg < G; gg++)
      if (sval[gg] > best) {
        best = sval[gg];
        d = sdv[gg];
        q = sidx[gg];
      }
    if (t == 0) piv[k] = q;
    const int gq = q / rpg, sq = q % rpg;
    if (act && g == gq) {
      // slot sq by selects (no per-slot branch), one division
      double y = 0;
#pragma unroll
      for (int k2 = 0; k2 < GJR_R; k2++) y = k2 == sq ? a[k2] : y;
      rowb[c] = (c == k ? 1.0 : y) / d;
      if (q != k) {
        const double x = rowx[c];
#pragma unroll
        for (int k2 = 0; k2 < GJR_R; k2++) a[k2] = k2 == sq ? x : a[k2];
      }
    }
    __syncthreads();
    // (C) row k from rowb; elimination of column k from every other row
    if (act) {
      const double rv = rowb[c];
      if (g == gk)
#pragma unroll
        for (int k2 = 0; k2 < GJR_R; k2++) a[k2] = k2 == sk ? rv : a[k2];
#pragma unroll
      for (int k2 = 0; k2 < GJR_R; k2++) {
        const int r = g * rpg + k2;
        if (k2 < rpg && r < n && r != k) {
          const double f = r == q ? colb[bf][k] : colb[bf][r];
          a[k2] = c == k ? 0.0 - f * rv : a[k2] - f * rv;
        }
      }
    }
  }
  __syncthreads();
  // undo the column swaps: output column pinv[c] = column c
  if (t == 0) {
    for (int j = 0; j < n; j++) pinv[j] = j;  // P
    for (int k = n - 1; k >= 0; k--) {
      const int q = piv[k], x = pinv[k];
      pinv[k] = pinv[q];
      pinv[q] = x;
    }
    // invert P through the (now free) pivot exchange buffer
    for (int j = 0; j < n; j++) pv[pinv[j]] = (double)j;
    for (int j = 0; j < n; j++) pinv[j] = (int)pv[j];
  }
  __syncthreads();
  if (act) {
    const int oc = pinv[c];
#pragma unroll
    for (int k2 = 0; k2 < GJR_R; k2++) {
      const int r = g * rpg + k2;
      if (k2 < rpg && r < n) inv[(size_t)r * n + oc] = a[k2];
    }
  }
}

static void gauss_jordan(int n, double *aug, double *inv, hipStream_t s);

// Gauss-Jordan inversion with partial pivoting in one workgroup (global memory)
template <class T>
__global__ void k_gauss_jordan(int n, double *a, T *inv) {
  __shared__ double sv[1024];
  __shared__ int si[1024];
  __shared__ int piv;
  const int t = threadIdx.x, nt = blockDim.x, w = 2 * n;
  for (int p = 0; p < n; p++) {
    double best = -1;
    int bi = p;
    for (int r = p + t; r < n; r += nt) {
      double v = fabs(a[(size_t)r * w + p]);
      if (v > best) {
        best = v;
        bi = r;
      }
    }
    sv[t] = best;
    si[t] = bi;
    __syncthreads();
    for (int o = nt / 2; o > 0; o >>= 1) {
      if (t < o && sv[t + o] > sv[t]) {
        sv[t] = sv[t + o];
        si[t] = si[t + o];
      }
      __syncthreads();
    }
    if (t == 0) piv = si[0];
    __syncthreads();
    const int q = piv;
    if (q != p)
      for (int c = t; c < w; c += nt) {
        double x = a[(size_t)p * w + c];
        a[(size_t)p * w + c] = a[(size_t)q * w + c];
        a[(size_t)q * w + c] = x;
      }
    __syncthreads();
    double d = a[(size_t)p * w + p];
    __syncthreads();
    for (int c = t; c < w; c += nt) a[(size_t)p * w + c] /= d;
    __syncthreads();
    for (int idx = t; idx < n * w; idx += nt) {
      int r = idx / w, c = idx % w;
      if (r == p) continue;
      double f = a[(size_t)r * w + p];
      if (c != p) a[(size_t)r * w + c] -= f * a[(size_t)p * w + c];
    }
    __syncthreads();
    for (int r = t; r < n; r += nt)
      if (r != p) a[(size_t)r * w + p] = 0.0;
    __syncthreads();
  }
  for (int idx = t; idx < n * n; idx += nt) inv[idx] = T(a[(size_t)(idx / n) * w + n + idx % n]);
}

static bool gj_reg_fits(int n) {
  if (n < 1 || n > GJR_N || GJR_T / n < 1) return false;
  const int G = GJR_T / n;
  return (n + G - 1) / G <= GJR_R;
}

static void gauss_jordan(int n, double *aug, double *inv, hipStream_t s) {
  if (n <= GJ_LDS_N)
    hipLaunchKernelGGL(k_gauss_jordan_lds, dim3(1), dim3(1024), 0, s, n, (const double *)aug, inv);
  else if (gj_reg_fits(n))
    hipLaunchKernelGGL(k_gauss_jordan_reg, dim3(1), dim3(GJR_T), 0, s, n, (const double *)aug, inv);
  else
    hipLaunchKernelGGL(k_gauss_jordan<double>, dim3(1), dim3(1024), 0, s, n, aug, inv);
}

// b = a^T (n x n)
__global__ void k_transpose(int n, const double *a, double *b) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (c < n) b[(size_t)c * n + r] = a[(size_t)r * n + c];
}

// u = Ainv f with f = [f0 (n0 entries), f1]
template <class T>
__global__ void k_dense_solve(int n, int n0, const double *ainv, const T *f0, const T *f1, T *u0, T *u1) {
  // one wavefront per row of the inverse: coalesced row reads, lane butterfly sum
  const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= n) return;
  double s = 0;
  for (int c = lane; c < n; c += 64) s += ainv[(size_t)r * n + c] * (double)(c < n0 ? f0[c] : f1[c - n0]);
  s = wave_sum(s);
  if (lane == 0) {
    if (r < n0)
      u0[r] = T(s);
    else
      u1[r - n0] = T(s);
  }
}


// ---------------------------------------------------------------- stencil levels
// Explicit stencils are stored slot-major: coefficient `slot` of row r
// (r = (comp-1)*N + m) at S[slot * 2N + r], so that the lanes of a wavefront
// (consecutive m) read consecutive addresses.  A window of half-width W
// holds 2*(2W+1)^2 slots: column component cc (0 Vx, 1 Vy) at offset (di,dj).
__device__ __forceinline__ int sslot(int W, int cc, int di, int dj) {
  const int w = 2 * W + 1;
  return cc * w * w + (di + W) * w + (dj + W);
}

// a*b+c with one rounding (the smoothers' row sums)
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class T>
struct StView {
  const T *S;
  int X, Y, W;
  const int8_t *kx, *ky;
};

// finest-level coefficients by probing the matrix-free rows with unit vectors
template <class T>
__global__ void k_probe0(VLev<T> L, T *S) {
  IDX2;
  const int X = L.o.X, Y = L.o.Y, N = X * Y;
  if (i >= X || j >= Y) return;
  const int m = i * Y + j;
  const size_t R = 2 * (size_t)N;
  for (int comp = 1; comp <= 2; comp++) {
    const size_t r = (size_t)(comp - 1) * N + m;
    for (int cc = 0; cc < 2; cc++)
      for (int di = -1; di <= 1; di++)
        for (int dj = -1; dj <= 1; dj++) {
          const int a = i + di, b = j + dj;
          T v = 0;
          if (a >= 0 && a < X && b >= 0 && b < Y) v = vrow(L, VUnit<T>{cc + 1, a * Y + b, Y}, comp, i, j);
          S[(size_t)sslot(1, cc, di, dj) * R + r] = v;
        }
  }
}

// Finest 3x3 stencil in difference form for FP32 iterates: row r reads
//   sum_cc [ sum_{k != centre} s_k (u_k - u_centre) + sigma_cc u_centre ]
// with sigma_cc the FP64 sum of the row's cc-component coefficients, stored
// in the centre slot.  Rounding the coefficients to FP32 then perturbs the
// couplings by 1e-7 relative but keeps the near-cancellation of the momentum
// rows on smooth fields (which plain FP32 coefficients lose).
template <class T>
__global__ void k_sdiff(const double *S, int N, T *D) {
  const size_t R = 2 * (size_t)N;
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  for (int cc = 0; cc < 2; cc++) {
    double sig = 0;
    for (int di = -1; di <= 1; di++)
      for (int dj = -1; dj <= 1; dj++) {
        const double v = S[(size_t)sslot(1, cc, di, dj) * R + r];
        sig += v;
        if (di != 0 || dj != 0) D[(size_t)sslot(1, cc, di, dj) * R + r] = T(v);
      }
    D[(size_t)sslot(1, cc, 0, 0) * R + r] = T(sig);
  }
}

// Galerkin coarse stencil R A P (R = normalised transpose of the bilinear
// prolongation restricted to equation rows, as the residual restriction is);
// pointwise rows (boundary, ghost) keep their mukoef-scaled form
// one thread per coarse point, 64 threads per block: the 50 accumulators of
// a row live in LDS (runtime slot indices would otherwise go to scratch)
// Galerkin product R A P of one coarse level: a block holds GAL_R coarse rows
// with GAL_S threads each, one per fine row of the row's restriction support
// (at most RW x RW).  Each thread accumulates its fine row's contributions in
// its own LDS slots; the GAL_S partial rows are then summed in a fixed order
// (reproducible).  Coarse pointwise rows get mu on the diagonal (and -mu k on
// their neighbour), as on the finest level.
constexpr int GAL_R = 4, GAL_S = RW * RW, GAL_B = GAL_R * GAL_S;
template <class T>
__global__ __launch_bounds__(GAL_B) void k_galerkin(StView<T> F, int Xc, int Yc, MapView xn, MapView xm, MapView yn,
                                                    MapView ym, const int8_t *ckx, const int8_t *cky, const int *cnbx,
                                                    const int *cnby, const T *ckkx, const T *ckky, double mu, T *Sc,
                                                    int *overflow) {
  __shared__ double sacc[50 * GAL_B];
  __shared__ double sw[GAL_B];
  const int r = threadIdx.x / GAL_S, sub = threadIdx.x % GAL_S;
  const int j = blockIdx.x * GAL_R + r, i = blockIdx.y;
  const bool rowok = i < Xc && j < Yc;
  double *acc = sacc + threadIdx.x;  // slot q at acc[q * GAL_B]
  const int Nc = Xc * Yc, Nf = F.X * F.Y, m = rowok ? i * Yc + j : 0;
  const size_t Rc = 2 * (size_t)Nc, Rf = 2 * (size_t)Nf;
  for (int comp = 1; comp <= 2; comp++) {
    const size_t rc = (size_t)(comp - 1) * Nc + m;
    for (int q = 0; q < 50; q++) acc[q * GAL_B] = 0.0;
    double wpart = 0;
    const int kind = rowok ? (comp == 1 ? ckx[m] : cky[m]) : -1;
    if (kind == 0) {
      const MapView mxr = comp == 1 ? xn : xm, myr = comp == 1 ? ym : yn;
      const int k = mxr.rb[i] + sub / RW, l = myr.rb[j] + sub % RW;
      const double w = (k < mxr.re[i] && l < myr.re[j]) ? wgt(mxr, k, i) * wgt(myr, l, j) : 0.0;
      if (w != 0) {
        wpart = w;
        const int a = k * F.Y + l;
        if ((comp == 1 ? F.kx[a] : F.ky[a]) == 0) {
          const size_t rf = (size_t)(comp - 1) * Nf + a;
          for (int cc = 0; cc < 2; cc++) {
            const MapView px = cc == 0 ? xn : xm, py = cc == 0 ? ym : yn;
            for (int dk = -F.W; dk <= F.W; dk++)
              for (int dl = -F.W; dl <= F.W; dl++) {
                const double sv = (double)F.S[(size_t)sslot(F.W, cc, dk, dl) * Rf + rf];
                if (sv == 0) continue;
                const int kb = k + dk, lb = l + dl;
                if (kb < 0 || kb >= F.X || lb < 0 || lb >= F.Y) continue;
                if (cc == 0 ? (lb == F.Y - 1) : (kb == F.X - 1)) continue;  // ghost: no prolongation
                const int a0 = px.i0[kb], b0 = py.i0[lb];
                const double wa = px.w[kb], wb = py.w[lb];
                for (int ia = 0; ia < 2; ia++)
                  for (int ib = 0; ib < 2; ib++) {
                    const double pw = (ia ? wa : 1.0 - wa) * (ib ? wb : 1.0 - wb);
                    if (pw == 0) continue;
                    const int di = a0 + ia - i, dj = b0 + ib - j;
                    if (di < -2 || di > 2 || dj < -2 || dj > 2) {
                      atomicAdd(overflow, 1);
                      continue;
                    }
                    acc[sslot(2, cc, di, dj) * GAL_B] += w * sv * pw;
                  }
              }
          }
        }
      }
    }
    sw[threadIdx.x] = wpart;
    __syncthreads();
    if (rowok) {
      const int base = r * GAL_S;
      if (kind != 0) {
        // pointwise row: mu on the diagonal, -mu k on the one neighbour
        const int nb = comp == 1 ? cnbx[m] : cnby[m];
        int qn = -1;
        double kk = 0;
        if (kind == 1 && nb >= 0) {
          const int di = nb / Yc - i, dj = nb % Yc - j;
          if (di >= -2 && di <= 2 && dj >= -2 && dj <= 2) {
            qn = sslot(2, comp - 1, di, dj);
            kk = (double)(comp == 1 ? ckkx[m] : ckky[m]);
          } else if (sub == 0) {
            atomicAdd(overflow, 1);
          }
        }
        const int qd = sslot(2, comp - 1, 0, 0);
        for (int q = sub; q < 50; q += GAL_S) {
          double v = (q == qd ? mu : 0.0);
          if (q == qn) v -= mu * kk;
          Sc[(size_t)q * Rc + rc] = T(v);
        }
      } else {
        double wsum = 0;
        for (int t = 0; t < GAL_S; t++) wsum += sw[base + t];
        const double inv = wsum > 0 ? 1.0 / wsum : 0.0;
        for (int q = sub; q < 50; q += GAL_S) {
          double v = 0;
          for (int t = 0; t < GAL_S; t++) v += sacc[q * GAL_B + base + t];
          Sc[(size_t)q * Rc + rc] = T(v * inv);
        }
      }
    }
    __syncthreads();
  }
}

// dense [Vx | Vy] matrix of the coarsest level from its stencil
template <class T>
__global__ void k_sdense(const T *S, int X, int Y, double *aug) {
  const int N = X * Y, n = 2 * N;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double *row = aug + (size_t)r * 2 * n;
  for (int c = 0; c < 2 * n; c++) row[c] = 0.0;
  row[n + r] = 1.0;
  const int m = r % N, i = m / Y, j = m % Y;
  for (int cc = 0; cc < 2; cc++)
    for (int di = -2; di <= 2; di++)
      for (int dj = -2; dj <= 2; dj++) {
        const int a = i + di, b = j + dj;
        if (a < 0 || a >= X || b < 0 || b >= Y) continue;
        row[cc * N + a * Y + b] += (double)S[(size_t)sslot(2, cc, di, dj) * n + r];
      }
}

// coarse-level residual r = f - S u on equation rows (0 on pointwise rows).
// Columns outside the grid carry zero coefficients, so clamped reads keep the
// 50-term sum free of branches.
template <class T>
__global__ void k_sresid(const T *S, int X, int Y, const int8_t *kx, const int8_t *ky, const T *ux, const T *uy,
                         const T *fx, const T *fy, T *rx, T *ry) {
  IDX2;
  if (i >= X || j >= Y) return;
  const int N = X * Y, m = i * Y + j;
  const size_t R = 2 * (size_t)N;
  int ia[5], jb[5];
#pragma unroll
  for (int d = 0; d < 5; d++) {
    const int a = i + d - 2, b = j + d - 2;
    ia[d] = (a < 0 ? 0 : (a >= X ? X - 1 : a)) * Y;
    jb[d] = b < 0 ? 0 : (b >= Y ? Y - 1 : b);
  }
  for (int comp = 0; comp < 2; comp++) {
    const T *Sr = S + (size_t)comp * N + m;
    T acc = 0;
    for (int cc = 0; cc < 2; cc++) {
      const T *u = cc ? uy : ux;
#pragma unroll
      for (int di = 0; di < 5; di++)
#pragma unroll
        for (int dj = 0; dj < 5; dj++) acc += Sr[(size_t)sslot(2, cc, di - 2, dj - 2) * R] * u[ia[di] + jb[dj]];
    }
    if (comp == 0)
      rx[m] = kx[m] == 0 ? fx[m] - acc : T(0);
    else
      ry[m] = ky[m] == 0 ? fy[m] - acc : T(0);
  }
}

// ---------------------------------------------------------------- tiled smoothers
// One workgroup smooths a TSY x TSX tile of a level in LDS (32 points along
// the contiguous j direction, so that each wavefront reads whole 128-byte
// lines of the stencil arrays): the tile and a halo
// of TH points are loaded once (with the coarse-grid correction added on the
// way up), nsmooth sweeps of the four colour steps (Vx red, Vx black, Vy red,
// Vy black; reversed after the coarse-grid correction) run between workgroup
// barriers, the pointwise boundary rows are re-solved, and the tile interior
// goes to the output buffers.  Halo values keep their kernel-start state
// (block Gauss-Seidel across tiles): no workgroup reads what another writes,
// so the result is deterministic.  Within a colour step every point of the
// colour is updated from the values before the step (on the Galerkin levels
// same-colour points couple; on the finest level they do not, and this is
// red-black Gauss-Seidel).  Every level smooths with its explicit stencil
// (finest: 3x3 window probed from the exact rows; coarser: 5x5 Galerkin).
constexpr int TH = 2;
// tile shape per precision: FP32 levels read whole 128-byte lines with 32
// points along j (and 32x16 tiles cut fewer couplings than 16x16: 88 vs 90
// Krylov iterations at 501^2); FP64 keeps 16x16, where the twice heavier
// colour steps of the larger tile cost more on the few-tile levels than the
// wider reads save
template <class T>
struct TileShape {
  static constexpr int SX = sizeof(T) == 4 ? 32 : 16, SY = 16, WX = SX + 2 * TH, WY = SY + 2 * TH, B = SX * SY;
};
#define TILE_CONSTS(T)                                                                                   \
  constexpr int TSX = TileShape<T>::SX, TSY = TileShape<T>::SY, TWX = TileShape<T>::WX, TWY = TileShape<T>::WY, \
                TB = TileShape<T>::B;                                                                     \
  (void)TSX; (void)TSY; (void)TWX; (void)TWY; (void)TB
#define TILE_CONSTS_XY(SX_, SY_)                                                                          \
  constexpr int TSX = SX_, TSY = SY_, TWX = SX_ + 2 * TH, TWY = SY_ + 2 * TH, TB = SX_ * SY_;            \
  (void)TSX; (void)TSY; (void)TWX; (void)TWY; (void)TB

// Tile shapes the autotuner chooses among for the Galerkin-level smoothing
// pass (VelMG::tune); candidate 0 is TileShape's.  Every shape keeps whole
// 128-byte lines along j for FP32 (32 or 64 points) or FP64 (16 or 32).
constexpr int kShapes = 4;
template <class T, int K>
struct ShapeCand;
template <> struct ShapeCand<float, 0> { static constexpr int SX = 32, SY = 16; };
template <> struct ShapeCand<float, 1> { static constexpr int SX = 32, SY = 8; };
template <> struct ShapeCand<float, 2> { static constexpr int SX = 64, SY = 8; };
template <> struct ShapeCand<float, 3> { static constexpr int SX = 32, SY = 32; };
template <> struct ShapeCand<double, 0> { static constexpr int SX = 16, SY = 16; };
template <> struct ShapeCand<double, 1> { static constexpr int SX = 16, SY = 8; };
template <> struct ShapeCand<double, 2> { static constexpr int SX = 32, SY = 8; };
template <> struct ShapeCand<double, 3> { static constexpr int SX = 32, SY = 16; };

template <class T>
struct TileArgs {
  int X, Y;
  const int8_t *kx, *ky;
  const T *fx, *fy;  // right-hand side
  const T *ux, *uy;  // input iterate (not read when zero != 0)
  T *ox, *oy;        // output iterate
  // correction on the next coarser level, added on load (ex == nullptr: none)
  const T *ex, *ey;
  int Yc;
  MapView xn, xm, yn, ym;  // this level -> next coarser
  int zero, nsmooth, rev;
  const T *dx, *dy;  // row diagonals (read by the difference-form stencil only)
  // pre-smoothing on the way down (k_tsmooth, rfx != nullptr): the
  // right-hand side is the restriction of the finer level's residual, formed
  // here (k_vrestrict's weights) and also stored to fxo, fyo for the later
  // residual and post-smoothing passes
  const T *rfx = nullptr, *rfy = nullptr;
  T *fxo = nullptr, *fyo = nullptr;
  int Yf = 0;
  MapView fxn{}, fxm{}, fyn{}, fym{};  // finer level -> this one
  double omega = 1.0;  // damping of the colour steps
};

template <class T>
__device__ __forceinline__ double prolong_at(const T *e, int Yc, const MapView &mx, const MapView &my, int i, int j) {
  const int a = mx.i0[i], b = my.i0[j];
  const double wa = mx.w[i], wb = my.w[j];
  return (1 - wa) * ((1 - wb) * (double)e[a * Yc + b] + wb * (double)e[a * Yc + b + 1]) +
         wa * ((1 - wb) * (double)e[(a + 1) * Yc + b] + wb * (double)e[(a + 1) * Yc + b + 1]);
}

template <class T, int SX = TileShape<T>::SX, int SY = TileShape<T>::SY>
__device__ __forceinline__ void tile_load(const TileArgs<T> &A, T *lx, T *ly, int i0, int j0) {
  TILE_CONSTS_XY(SX, SY);
  for (int q = threadIdx.x; q < TWX * TWY; q += TB) {
    const int i = i0 + q / TWX, j = j0 + q % TWX;
    T vx = 0, vy = 0;
    if (i >= 0 && i < A.X && j >= 0 && j < A.Y) {
      const int m = i * A.Y + j;
      if (!A.zero) {
        vx = A.ux[m];
        vy = A.uy[m];
      }
      if (A.ex) {
        if (A.kx[m] != 2 && j <= A.Y - 2) vx += T(prolong_at(A.ex, A.Yc, A.xn, A.ym, i, j));
        if (A.ky[m] != 2 && i <= A.X - 2) vy += T(prolong_at(A.ey, A.Yc, A.xm, A.yn, i, j));
      }
    }
    lx[q] = vx;
    ly[q] = vy;
  }
}

// S u at a tile point from LDS (out-of-grid halo entries are 0, and so are
// the coefficients that would multiply them)
template <class T, class TA, int W, bool DIFF>
__device__ __forceinline__ TA s_apply_lds(const TA *S, size_t R, size_t r, const T *lx, const T *ly, int la, int lb) {
  TILE_CONSTS(T);
  TA acc = 0;
#pragma unroll
  for (int cc = 0; cc < 2; cc++) {
    const T *u = cc ? ly : lx;
    const TA uc = DIFF ? TA(u[la * TWX + lb]) : TA(0);
#pragma unroll
    for (int di = -W; di <= W; di++)
#pragma unroll
      for (int dj = -W; dj <= W; dj++) {
        const TA s = S[(size_t)sslot(W, cc, di, dj) * R + r];
        const TA v = TA(u[(la + di) * TWX + lb + dj]);
        acc = fmadd(s, DIFF && (di != 0 || dj != 0) ? v - uc : v, acc);
      }
  }
  return acc;
}

// TA: type of the stencil and of the row sums (FP64 stencil with FP32
// iterates keeps the finest level's cancellations exact in mixed mode)
// DIFF: the finest level's FP32 stencil in difference form (k_sdiff), with
// the row diagonals read from A.dx / A.dy
// RS: pre-smoothing pass that restricts its right-hand side (A.rfx); a
// template flag, since the restriction code slows the plain pass's schedule
// even when not taken (13.8 -> 16.0 us for the FP32 Galerkin pass)
// SX, SY: the tile shape (TileShape's unless the autotuner chose another)
template <class T, class TA, int W, bool DIFF = false, bool RS = false, int SX = TileShape<T>::SX,
          int SY = TileShape<T>::SY>
__global__ __launch_bounds__(SX * SY) void k_tsmooth(TileArgs<T> A, const TA *S) {
  TILE_CONSTS_XY(SX, SY);
  __shared__ T lx[TWX * TWY], ly[TWX * TWY];
  const int i0 = blockIdx.y * TSY - TH, j0 = blockIdx.x * TSX - TH;
  const int X = A.X, Y = A.Y, N = X * Y;
  const size_t R = 2 * (size_t)N;
  tile_load<T, SX, SY>(A, lx, ly, i0, j0);
  const int i = i0 + TH + (int)threadIdx.x / TSX, j = j0 + TH + (int)threadIdx.x % TSX;
  const int li = (i - i0) * TWX + (j - j0);
  const bool ok = i < X && j < Y;
  const int m = ok ? i * Y + j : 0;
  const int kx = ok ? A.kx[m] : -1, ky = ok ? A.ky[m] : -1;
  T fx, fy;
  if constexpr (RS) {
    fx = kx == 0 && j <= Y - 2 ? restrict_rw(A.rfx, A.Yf, A.fxn, A.fym, i, j) : T(0);
    fy = ky == 0 && i <= X - 2 ? restrict_rw(A.rfy, A.Yf, A.fxm, A.fyn, i, j) : T(0);
    if (ok) {
      A.fxo[m] = fx;
      A.fyo[m] = fy;
    }
  } else {
    fx = ok ? A.fx[m] : T(0);
    fy = ok ? A.fy[m] : T(0);
  }
  // both stencil rows of this thread's point stay in registers for the pass
  constexpr int NSL = 2 * (2 * W + 1) * (2 * W + 1);
  TA sx[NSL], sy[NSL];
#pragma unroll
  for (int q = 0; q < NSL; q++) {
    sx[q] = ok ? S[(size_t)q * R + m] : TA(0);
    sy[q] = ok ? S[(size_t)q * R + (size_t)N + m] : TA(0);
  }
  const TA dx = ok ? (DIFF ? TA(A.dx[m]) : sx[sslot(W, 0, 0, 0)]) : TA(1);
  const TA dy = ok ? (DIFF ? TA(A.dy[m]) : sy[sslot(W, 1, 0, 0)]) : TA(1);
  const TA rdx = TA(A.omega) / dx, rdy = TA(A.omega) / dy;
  const int la = li / TWX, lb = li % TWX;
  auto apply = [&](const TA *sv) {
    TA acc = 0;
#pragma unroll
    for (int cc = 0; cc < 2; cc++) {
      const T *u = cc ? ly : lx;
      const TA uc = DIFF ? TA(u[la * TWX + lb]) : TA(0);
#pragma unroll
      for (int di = -W; di <= W; di++)
#pragma unroll
        for (int dj = -W; dj <= W; dj++) {
          const TA v = TA(u[(la + di) * TWX + lb + dj]);
          acc = fmadd(sv[sslot(W, cc, di, dj)], DIFF && (di != 0 || dj != 0) ? v - uc : v, acc);
        }
    }
    return acc;
  };
  // two colour steps on one component (compile-time, so that no per-slot
  // select between the Vx and Vy rows is needed)
  auto half = [&](auto comp, int c0) {
    constexpr int C = decltype(comp)::value;
    T *u = C == 1 ? lx : ly;
    const TA *sv = C == 1 ? sx : sy;
    const T f = C == 1 ? fx : fy;
    const TA rd = C == 1 ? rdx : rdy;
    for (int c = 0; c < 2; c++) {
      const bool act = ((i + j) & 1) == (c0 ^ c);
      T nv = 0;
      if (act) nv = T(u[li] + (f - apply(sv)) * rd);
      __syncthreads();
      if (act && ok) u[li] = nv;
      __syncthreads();
    }
  };
  __syncthreads();
  // Colour steps Vx red, Vx black, Vy red, Vy black (reversed after the
  // coarse-grid correction).  Lanes outside the grid run them too (zero
  // stencil, unit diagonal, in-tile LDS reads) and only skip the store: a
  // tile cut by the grid's last column otherwise leaves a sparse lane mask in
  // every wave, which measured 2x slower per sweep than the full checkerboard.
  for (int it = 0; it < A.nsmooth; it++) {
    if (!A.rev) {
      half(std::integral_constant<int, 1>{}, 0);
      half(std::integral_constant<int, 2>{}, 0);
    } else {
      half(std::integral_constant<int, 2>{}, 1);
      half(std::integral_constant<int, 1>{}, 1);
    }
  }
  if (!ok) return;
  // re-solve the pointwise rows exactly: their scale (mukoef) dwarfs the
  // momentum rows, and the last colour steps left them inconsistent
  T bx = lx[li], by = ly[li];
  if (kx > 0) bx = T(bx + (fx - apply(sx)) / dx);
  if (ky > 0) by = T(by + (fy - apply(sy)) / dy);
  A.ox[m] = bx;
  A.oy[m] = by;
}

// unknowns of the finest level read from the LDS tile (global indices)
template <class T>
struct LdsV {
  static constexpr int TWX = TileShape<T>::WX;
  const T *lx, *ly;
  int i0, j0;
  __device__ __forceinline__ T vx(int i, int j) const { return lx[(i - i0) * TWX + (j - j0)]; }
  __device__ __forceinline__ T vy(int i, int j) const { return ly[(i - i0) * TWX + (j - j0)]; }
  __device__ __forceinline__ T pp(int, int) const { return T(0); }
};

// finest level: the same tile scheme on the matrix-free rows (the nodal
// coefficient fields are read instead of an explicit 3x3 stencil, which
// would be several times the bytes), red-black Gauss-Seidel
template <class T>
__global__ __launch_bounds__(TileShape<T>::B) void k_tsmooth0(TileArgs<T> A, VLev<T> L, const T *dxa, const T *dya) {
  TILE_CONSTS(T);
  __shared__ T lx[TWX * TWY], ly[TWX * TWY];
  const int i0 = blockIdx.y * TSY - TH, j0 = blockIdx.x * TSX - TH;
  const int X = A.X, Y = A.Y;
  tile_load(A, lx, ly, i0, j0);
  const int i = i0 + TH + (int)threadIdx.x / TSX, j = j0 + TH + (int)threadIdx.x % TSX;
  const int li = (i - i0) * TWX + (j - j0);
  const bool ok = i < X && j < Y;
  const int m = ok ? i * Y + j : 0;
  const T fx = ok ? A.fx[m] : T(0), fy = ok ? A.fy[m] : T(0);
  const int kx = ok ? A.kx[m] : -1, ky = ok ? A.ky[m] : -1;
  const T dx = ok ? dxa[m] : T(1), dy = ok ? dya[m] : T(1);
  const LdsV<T> lv{lx, ly, i0, j0};
  __syncthreads();
  for (int it = 0; it < A.nsmooth; it++)
    for (int st = 0; st < 4; st++) {
      const int k = A.rev ? 3 - st : st, comp = 1 + k / 2, col = k & 1;
      const bool act = ok && ((i + j) & 1) == col;
      T nv = 0;
      if (act) nv = (comp == 1 ? lx[li] : ly[li]) + ((comp == 1 ? fx : fy) - vrow(L, lv, comp, i, j)) / (comp == 1 ? dx : dy);
      __syncthreads();
      if (act) (comp == 1 ? lx : ly)[li] = nv;
      __syncthreads();
    }
  if (!ok) return;
  // re-solve the pointwise rows exactly (their neighbour is adjacent, so in the tile)
  T bx = lx[li], by = ly[li];
  const T imu = T(1.0 / L.o.mu);
  if (kx > 0) {
    const int nb = L.nbx[m];
    bx = fx * imu + (kx == 1 && nb >= 0 ? L.kkx[m] * lv.vx(nb / Y, nb % Y) : T(0));
  }
  if (ky > 0) {
    const int nb = L.nby[m];
    by = fy * imu + (ky == 1 && nb >= 0 ? L.kky[m] * lv.vy(nb / Y, nb % Y) : T(0));
  }
  A.ox[m] = bx;
  A.oy[m] = by;
}

// r = f - S u on equation rows, 0 on pointwise rows (u staged through LDS)
template <class T, class TA, int W, bool DIFF = false>
__global__ __launch_bounds__(TileShape<T>::B) void k_tresid(TileArgs<T> A, const TA *S, T *rxo, T *ryo) {
  TILE_CONSTS(T);
  __shared__ T lx[TWX * TWY], ly[TWX * TWY];
  const int i0 = blockIdx.y * TSY - TH, j0 = blockIdx.x * TSX - TH;
  const int X = A.X, Y = A.Y, N = X * Y;
  const size_t R = 2 * (size_t)N;
  tile_load(A, lx, ly, i0, j0);
  __syncthreads();
  const int i = i0 + TH + (int)threadIdx.x / TSX, j = j0 + TH + (int)threadIdx.x % TSX;
  if (i >= X || j >= Y) return;
  const int m = i * Y + j, la = i - i0, lb = j - j0;
  rxo[m] = A.kx[m] == 0 ? T(A.fx[m] - s_apply_lds<T, TA, W, DIFF>(S, R, m, lx, ly, la, lb)) : T(0);
  ryo[m] = A.ky[m] == 0 ? T(A.fy[m] - s_apply_lds<T, TA, W, DIFF>(S, R, (size_t)N + m, lx, ly, la, lb)) : T(0);
}

// The two coarsest velocity levels in one workgroup: ns colour sweeps on the
// second-coarsest level (the whole level as one tile in LDS, zero outside
// the grid), its residual restricted to the coarsest level, the coarsest
// solve by the dense inverse, the prolongated correction and ns reversed
// sweeps, the pointwise rows re-solved after each smoothing.  Each part does
// the arithmetic of k_tsmooth, k_tresid and k_vrestrict; one tile instead of
// several only drops the frozen halo between tiles, and the dense product
// sums each row in column order.  Four
// launches per V-cycle fewer.
constexpr int VT_S = 17, VT_P = VT_S + 2 * TH, VT_B = 320, VT_C = 81;
template <class T>
struct VTail {
  int X, Y, Xc, Yc, ns;
  const int8_t *kx, *ky, *ckx, *cky;
  const T *S, *fx, *fy;
  MapView xn, xm, yn, ym;  // this level -> coarsest
  const double *ainvT;     // transposed inverse, (2 Xc Yc)^2
  T *ox, *oy;
  double omega;
};
template <class T>
__global__ __launch_bounds__(VT_B) void k_vtail(VTail<T> A) {
  __shared__ T lx[VT_P * VT_P], ly[VT_P * VT_P], rx[VT_S * VT_S], ry[VT_S * VT_S], fc[2 * VT_C], uc[2 * VT_C];
  const int t = threadIdx.x, X = A.X, Y = A.Y, N = X * Y;
  const size_t R = 2 * (size_t)N;
  for (int q = t; q < VT_P * VT_P; q += VT_B) {
    lx[q] = T(0);
    ly[q] = T(0);
  }
  // lanes past the level run the colour steps at (0,0) with a zero stencil
  // and never store
  const bool ok = t < N;
  const int i = ok ? t / Y : 0, j = ok ? t % Y : 0, m = i * Y + j;
  const int la = i + TH, lb = j + TH, li = la * VT_P + lb;
  const int kx = ok ? A.kx[m] : -1, ky = ok ? A.ky[m] : -1;
  const T fx = ok ? A.fx[m] : T(0), fy = ok ? A.fy[m] : T(0);
  constexpr int NSL = 50;
  T sx[NSL], sy[NSL];
#pragma unroll
  for (int q = 0; q < NSL; q++) {
    sx[q] = ok ? A.S[(size_t)q * R + m] : T(0);
    sy[q] = ok ? A.S[(size_t)q * R + (size_t)N + m] : T(0);
  }
  const T dx = ok ? sx[sslot(2, 0, 0, 0)] : T(1), dy = ok ? sy[sslot(2, 1, 0, 0)] : T(1);
  const T rdx = T(A.omega) / dx, rdy = T(A.omega) / dy;
  auto apply = [&](const T *sv) {
    T acc = 0;
#pragma unroll
    for (int cc = 0; cc < 2; cc++) {
      const T *u = cc ? ly : lx;
#pragma unroll
      for (int di = -2; di <= 2; di++)
#pragma unroll
        for (int dj = -2; dj <= 2; dj++) acc = fmadd(sv[sslot(2, cc, di, dj)], u[(la + di) * VT_P + lb + dj], acc);
    }
    return acc;
  };
  auto half = [&](auto comp, int c0) {
    constexpr int C = decltype(comp)::value;
    T *u = C == 1 ? lx : ly;
    const T *sv = C == 1 ? sx : sy;
    const T f = C == 1 ? fx : fy;
    const T rd = C == 1 ? rdx : rdy;
    for (int c = 0; c < 2; c++) {
      const bool act = ((i + j) & 1) == (c0 ^ c);
      T nv = 0;
      if (act) nv = T(u[li] + (f - apply(sv)) * rd);
      __syncthreads();
      if (act && ok) u[li] = nv;
      __syncthreads();
    }
  };
  // pointwise rows re-solved exactly from the smoothed neighbours
  auto resolve = [&](T &bx, T &by) {
    bx = lx[li];
    by = ly[li];
    if (kx > 0) bx = T(bx + (fx - apply(sx)) / dx);
    if (ky > 0) by = T(by + (fy - apply(sy)) / dy);
  };
  __syncthreads();
  for (int it = 0; it < A.ns; it++) {
    half(std::integral_constant<int, 1>{}, 0);
    half(std::integral_constant<int, 2>{}, 0);
  }
  T bx, by;
  resolve(bx, by);
  __syncthreads();
  if (ok) {
    lx[li] = bx;
    ly[li] = by;
  }
  __syncthreads();
  if (ok) {
    rx[m] = kx == 0 ? T(fx - apply(sx)) : T(0);
    ry[m] = ky == 0 ? T(fy - apply(sy)) : T(0);
  }
  __syncthreads();
  const int Nc = A.Xc * A.Yc, n = 2 * Nc;
  for (int q = t; q < Nc; q += VT_B) {
    const int ic = q / A.Yc, jc = q % A.Yc;
    fc[q] = (A.ckx[q] == 0 && jc <= A.Yc - 2) ? restrict_rw((const T *)rx, Y, A.xn, A.ym, ic, jc) : T(0);
    fc[Nc + q] = (A.cky[q] == 0 && ic <= A.Xc - 2) ? restrict_rw((const T *)ry, Y, A.xm, A.yn, ic, jc) : T(0);
  }
  __syncthreads();
  // one thread per row, columns of the transposed inverse (coalesced, the
  // loads independent of the running sum)
  for (int r = t; r < n; r += VT_B) {
    double sum = 0;
#pragma unroll 8
    for (int c = 0; c < n; c++) sum += A.ainvT[(size_t)c * n + r] * (double)fc[c];
    uc[r] = T(sum);
  }
  __syncthreads();
  if (ok) {
    T vx = lx[li], vy = ly[li];
    if (kx != 2 && j <= Y - 2) vx += T(prolong_at((const T *)uc, A.Yc, A.xn, A.ym, i, j));
    if (ky != 2 && i <= X - 2) vy += T(prolong_at((const T *)(uc + Nc), A.Yc, A.xm, A.yn, i, j));
    lx[li] = vx;
    ly[li] = vy;
  }
  __syncthreads();
  for (int it = 0; it < A.ns; it++) {
    half(std::integral_constant<int, 2>{}, 1);
    half(std::integral_constant<int, 1>{}, 1);
  }
  resolve(bx, by);
  if (!ok) return;
  A.ox[m] = bx;
  A.oy[m] = by;
}

// ---------------------------------------------------------------- coarse V-cycle, one launch
// The FP32 levels from lc down to the two-level tail in one cooperative
// launch: every workgroup owns one 32x16 tile of one level (or, the last
// one, the tail), loads that tile's two stencil rows per point into
// registers once, and keeps them for the pre-smoothing pass, the residual
// and the post-smoothing pass; the phases (pre-smoothing and residual per
// level downwards, the tail, post-smoothing upwards) are separated by
// grid-wide barriers, which every workgroup reaches.  Each phase does the
// arithmetic of k_tsmooth (the restricting variant for every level after
// the first), k_tresid and k_vtail on the same tiles, so the V-cycle equals
// the launch sequence's bit for bit; it replaces 4 launches per level and
// the tail's with one, and the stencils are read once instead of three
// times.  The arguments (several KB) live in device memory.
constexpr int VC_MAXL = 4, VC_B = 512;
template <class T>
struct VCoarse {
  int nl;                    // tiled levels before the tail
  int b0[VC_MAXL + 1];       // first block of each level; b0[nl]: the tail's block
  int ntj[VC_MAXL];          // tiles along j of each level
  // pre-smoothing (from zero; restricting its right-hand side after the
  // first level), residual of its output, post-smoothing
  TileArgs<T> pre[VC_MAXL], res[VC_MAXL], post[VC_MAXL];
  const T *S[VC_MAXL];
  T *rx[VC_MAXL], *ry[VC_MAXL];  // residual of each level's pre-smoothed iterate
  VTail<T> tail;                 // its fx, fy: restricted here from rx[nl - 1]
  T *tfx, *tfy;
  int Yf;                        // Y of level nl - 1
  MapView xn, xm, yn, ym;        // level nl - 1 -> the tail level
};

template <class T>
__global__ __launch_bounds__(VC_B) void k_vcoarse(const VCoarse<T> *__restrict__ Vp) {
  namespace cg = cooperative_groups;
  const VCoarse<T> &V = *Vp;
  cg::grid_group grid = cg::this_grid();
  constexpr int SX = 32, SY = 16, TWX = SX + 2 * TH, TWY = SY + 2 * TH, NSL = 50;
  static_assert(SX * SY == VC_B, "one thread per tile point");
  __shared__ T lx[TWX * TWY], ly[TWX * TWY];
  __shared__ T trx[VT_S * VT_S], try_[VT_S * VT_S], tfc[2 * VT_C], tuc[2 * VT_C];
  const int b = blockIdx.x, tid = threadIdx.x;
  int lev = V.nl;  // this block's level (nl: the tail)
  for (int k = 0; k < V.nl; k++)
    if (b >= V.b0[k] && b < V.b0[k + 1]) lev = k;
  const bool tiled = lev < V.nl;
  // the tile of this block, its point, and the point's stencil rows
  int i0 = 0, j0 = 0, i = 0, j = 0, li = 0, m = 0, kx = -1, ky = -1;
  bool ok = false;
  T sx[NSL], sy[NSL], fx = 0, fy = 0, dx = 1, dy = 1;
  if (tiled) {
    const TileArgs<T> &A = V.pre[lev];
    const int t = b - V.b0[lev];
    i0 = (t / V.ntj[lev]) * SY - TH;
    j0 = (t % V.ntj[lev]) * SX - TH;
    i = i0 + TH + tid / SX;
    j = j0 + TH + tid % SX;
    li = (i - i0) * TWX + (j - j0);
    ok = i < A.X && j < A.Y;
    m = ok ? i * A.Y + j : 0;
    kx = ok ? A.kx[m] : -1;
    ky = ok ? A.ky[m] : -1;
    const size_t R = 2 * (size_t)A.X * A.Y, N = (size_t)A.X * A.Y;
#pragma unroll
    for (int q = 0; q < NSL; q++) {
      sx[q] = ok ? V.S[lev][(size_t)q * R + m] : T(0);
      sy[q] = ok ? V.S[lev][(size_t)q * R + N + m] : T(0);
    }
    dx = ok ? sx[sslot(2, 0, 0, 0)] : T(1);
    dy = ok ? sy[sslot(2, 1, 0, 0)] : T(1);
  }
  const int la = i - i0, lb = j - j0;
  auto apply = [&](const T *sv) {
    T acc = 0;
#pragma unroll
    for (int cc = 0; cc < 2; cc++) {
      const T *u = cc ? ly : lx;
#pragma unroll
      for (int di = -2; di <= 2; di++)
#pragma unroll
        for (int dj = -2; dj <= 2; dj++) acc = fmadd(sv[sslot(2, cc, di, dj)], u[(la + di) * TWX + lb + dj], acc);
    }
    return acc;
  };
  // one smoothing pass of k_tsmooth<T, T, 2, false, RS, 32, 16> on this tile
  auto pass = [&](const TileArgs<T> &A, bool RS) {
    tile_load<T, SX, SY>(A, lx, ly, i0, j0);
    if (RS) {
      fx = kx == 0 && j <= A.Y - 2 ? restrict_rw(A.rfx, A.Yf, A.fxn, A.fym, i, j) : T(0);
      fy = ky == 0 && i <= A.X - 2 ? restrict_rw(A.rfy, A.Yf, A.fxm, A.fyn, i, j) : T(0);
      if (ok) {
        A.fxo[m] = fx;
        A.fyo[m] = fy;
      }
    } else {
      fx = ok ? A.fx[m] : T(0);
      fy = ok ? A.fy[m] : T(0);
    }
    const T rdx = T(A.omega) / dx, rdy = T(A.omega) / dy;
    auto half = [&](auto comp, int c0) {
      constexpr int C = decltype(comp)::value;
      T *u = C == 1 ? lx : ly;
      const T *sv = C == 1 ? sx : sy;
      const T f = C == 1 ? fx : fy;
      const T rd = C == 1 ? rdx : rdy;
      for (int c = 0; c < 2; c++) {
        const bool act = ((i + j) & 1) == (c0 ^ c);
        T nv = 0;
        if (act) nv = T(u[li] + (f - apply(sv)) * rd);
        __syncthreads();
        if (act && ok) u[li] = nv;
        __syncthreads();
      }
    };
    __syncthreads();
    for (int it = 0; it < A.nsmooth; it++) {
      if (!A.rev) {
        half(std::integral_constant<int, 1>{}, 0);
        half(std::integral_constant<int, 2>{}, 0);
      } else {
        half(std::integral_constant<int, 2>{}, 1);
        half(std::integral_constant<int, 1>{}, 1);
      }
    }
    if (ok) {
      T bx = lx[li], by = ly[li];
      if (kx > 0) bx = T(bx + (fx - apply(sx)) / dx);
      if (ky > 0) by = T(by + (fy - apply(sy)) / dy);
      A.ox[m] = bx;
      A.oy[m] = by;
    }
    __syncthreads();
  };
  // k_tresid<T, T, 2> on this tile: r = f - S u of the pre-smoothed iterate
  auto resid = [&](const TileArgs<T> &Rg, T *rxo, T *ryo) {
    tile_load<T, SX, SY>(Rg, lx, ly, i0, j0);
    __syncthreads();
    if (ok) {
      rxo[m] = kx == 0 ? T(fx - apply(sx)) : T(0);
      ryo[m] = ky == 0 ? T(fy - apply(sy)) : T(0);
    }
    __syncthreads();
  };
  // downwards: pre-smoothing and residual, level by level
  for (int k = 0; k < V.nl; k++) {
    if (lev == k) pass(V.pre[k], k > 0);
    grid.sync();
    if (lev == k) resid(V.res[k], V.rx[k], V.ry[k]);
    grid.sync();
  }
  // the tail (k_vtail), its right-hand side restricted from the last level
  if (!tiled) {
    const VTail<T> &A = V.tail;
    const int t = tid, X = A.X, Y = A.Y, N = X * Y;
    const size_t R = 2 * (size_t)N;
    constexpr int P = VT_P;
    for (int q = t; q < P * P; q += VC_B) {
      lx[q] = T(0);
      ly[q] = T(0);
    }
    const bool tok = t < N;
    const int ti = tok ? t / Y : 0, tj = tok ? t % Y : 0, tm = ti * Y + tj;
    const int ta = ti + TH, tb = tj + TH, tli = ta * P + tb;
    const int tkx = tok ? A.kx[tm] : -1, tky = tok ? A.ky[tm] : -1;
    // k_vrestrict into this level's right-hand side
    const T tfx = tok ? ((tkx == 0 && tj <= Y - 2) ? restrict_rw(V.rx[V.nl - 1], V.Yf, V.xn, V.ym, ti, tj) : T(0)) : T(0);
    const T tfy = tok ? ((tky == 0 && ti <= X - 2) ? restrict_rw(V.ry[V.nl - 1], V.Yf, V.xm, V.yn, ti, tj) : T(0)) : T(0);
    if (tok) {
      V.tfx[tm] = tfx;
      V.tfy[tm] = tfy;
    }
#pragma unroll
    for (int q = 0; q < NSL; q++) {
      sx[q] = tok ? A.S[(size_t)q * R + tm] : T(0);
      sy[q] = tok ? A.S[(size_t)q * R + (size_t)N + tm] : T(0);
    }
    const T tdx = tok ? sx[sslot(2, 0, 0, 0)] : T(1), tdy = tok ? sy[sslot(2, 1, 0, 0)] : T(1);
    const T rdx = T(A.omega) / tdx, rdy = T(A.omega) / tdy;
    auto tapply = [&](const T *sv) {
      T acc = 0;
#pragma unroll
      for (int cc = 0; cc < 2; cc++) {
        const T *u = cc ? ly : lx;
#pragma unroll
        for (int di = -2; di <= 2; di++)
#pragma unroll
          for (int dj = -2; dj <= 2; dj++) acc = fmadd(sv[sslot(2, cc, di, dj)], u[(ta + di) * P + tb + dj], acc);
      }
      return acc;
    };
    auto thalf = [&](auto comp, int c0) {
      constexpr int C = decltype(comp)::value;
      T *u = C == 1 ? lx : ly;
      const T *sv = C == 1 ? sx : sy;
      const T f = C == 1 ? tfx : tfy;
      const T rd = C == 1 ? rdx : rdy;
      for (int c = 0; c < 2; c++) {
        const bool act = ((ti + tj) & 1) == (c0 ^ c);
        T nv = 0;
        if (act) nv = T(u[tli] + (f - tapply(sv)) * rd);
        __syncthreads();
        if (act && tok) u[tli] = nv;
        __syncthreads();
      }
    };
    auto resolve = [&](T &bx, T &by) {
      bx = lx[tli];
      by = ly[tli];
      if (tkx > 0) bx = T(bx + (tfx - tapply(sx)) / tdx);
      if (tky > 0) by = T(by + (tfy - tapply(sy)) / tdy);
    };
    __syncthreads();
    for (int it = 0; it < A.ns; it++) {
      thalf(std::integral_constant<int, 1>{}, 0);
      thalf(std::integral_constant<int, 2>{}, 0);
    }
    T bx, by;
    resolve(bx, by);
    __syncthreads();
    if (tok) {
      lx[tli] = bx;
      ly[tli] = by;
    }
    __syncthreads();
    if (tok) {
      trx[tm] = tkx == 0 ? T(tfx - tapply(sx)) : T(0);
      try_[tm] = tky == 0 ? T(tfy - tapply(sy)) : T(0);
    }
    __syncthreads();
    const int Nc = A.Xc * A.Yc, n = 2 * Nc;
    for (int q = t; q < Nc; q += VC_B) {
      const int ic = q / A.Yc, jc = q % A.Yc;
      tfc[q] = (A.ckx[q] == 0 && jc <= A.Yc - 2) ? restrict_rw((const T *)trx, Y, A.xn, A.ym, ic, jc) : T(0);
      tfc[Nc + q] = (A.cky[q] == 0 && ic <= A.Xc - 2) ? restrict_rw((const T *)try_, Y, A.xm, A.yn, ic, jc) : T(0);
    }
    __syncthreads();
    for (int r = t; r < n; r += VC_B) {
      double sum = 0;
#pragma unroll 8
      for (int c = 0; c < n; c++) sum += A.ainvT[(size_t)c * n + r] * (double)tfc[c];
      tuc[r] = T(sum);
    }
    __syncthreads();
    if (tok) {
      T vx = lx[tli], vy = ly[tli];
      if (tkx != 2 && tj <= Y - 2) vx += T(prolong_at((const T *)tuc, A.Yc, A.xn, A.ym, ti, tj));
      if (tky != 2 && ti <= X - 2) vy += T(prolong_at((const T *)(tuc + Nc), A.Yc, A.xm, A.yn, ti, tj));
      lx[tli] = vx;
      ly[tli] = vy;
    }
    __syncthreads();
    for (int it = 0; it < A.ns; it++) {
      thalf(std::integral_constant<int, 2>{}, 1);
      thalf(std::integral_constant<int, 1>{}, 1);
    }
    resolve(bx, by);
    if (tok) {
      A.ox[tm] = bx;
      A.oy[tm] = by;
    }
  }
  grid.sync();
  // upwards: post-smoothing with the coarse correction added on load
  for (int k = V.nl - 1; k >= 0; k--) {
    if (lev == k) pass(V.post[k], false);
    if (k > 0) grid.sync();
  }
}

// ---------------------------------------------------------------- VelMG host
template <class T>
VLev<T> VelMG<T>::lev(int l) const {
  const VelLevel<T> &v = L[l];
  const GeoLevel &g = (*geo)[l];
  VLev<T> o;
  o.o = SOp<T>{v.X, v.Y, g.gx.p, g.gy.p, v.nu.p, v.nd.p, v.ro.p, v.gp.p, dt, mu};
  o.kx = v.kx.p;
  o.ky = v.ky.p;
  o.nbx = v.nbx.p;
  o.nby = v.nby.p;
  o.kkx = v.kkx.p;
  o.kky = v.kky.p;
  return o;
}

// BC pattern of a coarse level from the finer one: each coarse unknown copies
// kind and neighbour offset of the fine unknown on the same wall position
static int map_index(int I, int nc, int nf, const std::vector<int> &fi, int lastpad) {
  // lastpad: 1 for mid-point sets (last valid index nc-2 <-> nf-2)
  if (I == 0) return 0;
  if (I >= nc - 1 - lastpad) return nf - (nc - I);
  int a = fi[I];
  if (a < 1) a = 1;
  if (a > nf - 2 - lastpad) a = nf - 2 - lastpad;
  return a;
}

template <class T>
int VelMG<T>::setup(const std::vector<GeoLevel> &g, const std::vector<int8_t> &kx0, const std::vector<int8_t> &ky0,
                    const std::vector<int> &nbx0, const std::vector<int> &nby0, const std::vector<double> &kkx0,
                    const std::vector<double> &kky0) {
  shape.clear();
  vgen++;
  geo = &g;
  const int nl = (int)g.size();
  L.clear();
  L.resize(nl);
  for (int l = 0; l < nl; l++) {
    VelLevel<T> &v = L[l];
    v.X = g[l].X;
    v.Y = g[l].Y;
    v.N = v.X * v.Y;
    if (l == 0) {
      v.hkx = kx0;
      v.hky = ky0;
      v.hnbx = nbx0;
      v.hnby = nby0;
      v.hkkx.assign(kkx0.begin(), kkx0.end());
      v.hkky.assign(kky0.begin(), kky0.end());
    } else {
      const VelLevel<T> &f = L[l - 1];
      const GeoLevel &gf = g[l - 1];
      v.hkx.assign(v.N, 0);
      v.hky.assign(v.N, 0);
      v.hnbx.assign(v.N, -1);
      v.hnby.assign(v.N, -1);
      v.hkkx.assign(v.N, T(0));
      v.hkky.assign(v.N, T(0));
      for (int I = 0; I < v.X; I++)
        for (int J = 0; J < v.Y; J++) {
          int m = I * v.Y + J;
          // Vx: x nodes, y mids (ghost row J = Y-1)
          {
            int a = map_index(I, v.X, f.X, gf.fi, 0), b = (J == v.Y - 1) ? f.Y - 1 : map_index(J, v.Y, f.Y, gf.fj, 1);
            int fm = a * f.Y + b;
            v.hkx[m] = f.hkx[fm];
            if (f.hkx[fm] == 1 && f.hnbx[fm] >= 0) {
              int da = f.hnbx[fm] / f.Y - a, db = f.hnbx[fm] % f.Y - b;
              v.hnbx[m] = (I + da) * v.Y + (J + db);
              v.hkkx[m] = f.hkkx[fm];
            }
          }
          {
            int a = (I == v.X - 1) ? f.X - 1 : map_index(I, v.X, f.X, gf.fi, 1), b = map_index(J, v.Y, f.Y, gf.fj, 0);
            int fm = a * f.Y + b;
            v.hky[m] = f.hky[fm];
            if (f.hky[fm] == 1 && f.hnby[fm] >= 0) {
              int da = f.hnby[fm] / f.Y - a, db = f.hnby[fm] % f.Y - b;
              v.hnby[m] = (I + da) * v.Y + (J + db);
              v.hkky[m] = f.hkky[fm];
            }
          }
        }
    }
    ST_OK(v.nu.alloc(v.N));
    ST_OK(v.nd.alloc(v.N));
    ST_OK(v.ro.alloc(v.N));
    ST_OK(v.gp.alloc(v.N));
    ST_OK(v.kx.alloc(v.N));
    ST_OK(v.ky.alloc(v.N));
    ST_OK(v.nbx.alloc(v.N));
    ST_OK(v.nby.alloc(v.N));
    ST_OK(v.kkx.alloc(v.N));
    ST_OK(v.kky.alloc(v.N));
    ST_OK(v.dx.alloc(v.N));
    ST_OK(v.dy.alloc(v.N));
    ST_OK(v.ux.alloc(v.N));
    ST_OK(v.uy.alloc(v.N));
    ST_OK(v.fx.alloc(v.N));
    ST_OK(v.fy.alloc(v.N));
    ST_OK(v.rx.alloc(v.N));
    ST_OK(v.ry.alloc(v.N));
    ST_OK(v.ux2.alloc(v.N));
    ST_OK(v.uy2.alloc(v.N));
    v.W = l == 0 ? 1 : 2;
    ST_OK(v.S.alloc((size_t)2 * v.N * (l == 0 ? 18 : 50)));
    HIP_OK(hipMemcpy(v.kx.p, v.hkx.data(), v.N, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(v.ky.p, v.hky.data(), v.N, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(v.nbx.p, v.hnbx.data(), v.N * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(v.nby.p, v.hnby.data(), v.N * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(v.kkx.p, v.hkkx.data(), v.N * sizeof(T), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(v.kky.p, v.hkky.data(), v.N * sizeof(T), hipMemcpyHostToDevice));
  }
  gfi.resize(nl);
  gfj.resize(nl);
  for (int l = 0; l + 1 < nl; l++) {
    const GeoLevel &gf = g[l], &gc = g[l + 1];
    ST_OK(gfi[l].alloc(gf.fi.size()));
    ST_OK(gfj[l].alloc(gf.fj.size()));
    HIP_OK(hipMemcpy(gfi[l].p, gf.fi.data(), gf.fi.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(gfj[l].p, gf.fj.data(), gf.fj.size() * sizeof(int), hipMemcpyHostToDevice));
    auto mids = [](const std::vector<double> &a) {
      std::vector<double> m;
      for (size_t k = 0; k + 1 < a.size(); k++) m.push_back((a[k] + a[k + 1]) / 2.0);
      return m;
    };
    ST_OK(L[l].mxn.build(gf.hx, gc.hx));
    ST_OK(L[l].mxm.build(mids(gf.hx), mids(gc.hx)));
    ST_OK(L[l].myn.build(gf.hy, gc.hy));
    ST_OK(L[l].mym.build(mids(gf.hy), mids(gc.hy)));
  }
  nc = 2 * L[nl - 1].N;
  ST_OK(dense.alloc((size_t)nc * 2 * nc));
  ST_OK(ainv.alloc((size_t)nc * nc));
  ST_OK(ainvT.alloc((size_t)nc * nc));
  ST_OK(overflow.alloc(1));
  HIP_OK(hipMemset(overflow.p, 0, sizeof(int)));
  return I2E_OK;
}

template <class T>
int VelMG<T>::refresh(const double *nu, const double *nd, const double *ro, const double *gp, double mu_, double dt_,
                      hipStream_t s, hipStream_t side_) {
  ST_OK(join(s));  // a previous inversion still running (refresh without a join)
  vgen++;
  mu = mu_;
  dt = dt_;
  const int nl = (int)L.size();
  const int N0 = L[0].N;
  hipLaunchKernelGGL(k_cast<T>, dim3(nblk(N0, 256)), dim3(256), 0, s, N0, nu, L[0].nu.p);
  hipLaunchKernelGGL(k_cast<T>, dim3(nblk(N0, 256)), dim3(256), 0, s, N0, nd, L[0].nd.p);
  hipLaunchKernelGGL(k_cast<T>, dim3(nblk(N0, 256)), dim3(256), 0, s, N0, ro, L[0].ro.p);
  hipLaunchKernelGGL(k_cast<T>, dim3(nblk(N0, 256)), dim3(256), 0, s, N0, gp, L[0].gp.p);
  // finest level: diagonal for the matrix-free smoother and probed stencil
  hipLaunchKernelGGL(k_vdiag<T>, grd2(L[0].X, L[0].Y), blk2(), 0, s, lev(0), L[0].dx.p, L[0].dy.p);
  hipLaunchKernelGGL(k_probe0<T>, grd2(L[0].X, L[0].Y), blk2(), 0, s, lev(0), L[0].S.p);
  for (int l = 0; l + 1 < nl; l++) {
    VelLevel<T> &f = L[l], &c = L[l + 1];
    StView<T> fv{f.S.p, f.X, f.Y, f.W, f.kx.p, f.ky.p};
    hipLaunchKernelGGL(k_galerkin<T>, dim3((c.Y + GAL_R - 1) / GAL_R, c.X), dim3(GAL_B), 0, s, fv, c.X, c.Y,
                       view(f.mxn), view(f.mxm),
                       view(f.myn), view(f.mym), c.kx.p, c.ky.p, c.nbx.p, c.nby.p, c.kkx.p, c.kky.p, mu_, c.S.p,
                       overflow.p);
  }
  const int n = nc;
  if (nl > 1)
    hipLaunchKernelGGL(k_sdense<T>, dim3(nblk(n, 64)), dim3(64), 0, s, L[nl - 1].S.p, L[nl - 1].X, L[nl - 1].Y,
                       dense.p);
  else
    hipLaunchKernelGGL(k_vdense<T>, dim3(nblk(n, 64), n), dim3(64), 0, s, lev(0), n, dense.p);
  if (side_) {
    // the inversion (one workgroup, ~1 ms at 162 unknowns) overlaps the
    // caller's further set-up on s until join()
    if (!ev_fork) HIP_OK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    if (!ev_join) HIP_OK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    HIP_OK(hipEventRecord(ev_fork, s));
    HIP_OK(hipStreamWaitEvent(side_, ev_fork, 0));
    gauss_jordan(n, dense.p, ainv.p, side_);
    hipLaunchKernelGGL(k_transpose, dim3(nblk(n, 256), n), dim3(256), 0, side_, n, (const double *)ainv.p, ainvT.p);
    side = side_;
  } else {
    side = nullptr;
    gauss_jordan(n, dense.p, ainv.p, s);
    hipLaunchKernelGGL(k_transpose, dim3(nblk(n, 256), n), dim3(256), 0, s, n, (const double *)ainv.p, ainvT.p);
  }
  HIP_OK(hipGetLastError());
  int nover = 0;
  HIP_OK(hipMemcpyAsync(&nover, overflow.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (nover) {
    set_error("velocity multigrid: %d Galerkin couplings fall outside the 5x5 stencil window", nover);
    return I2E_ERR_UNSUPPORTED;
  }
  return I2E_OK;
}

template <class T>
int VelMG<T>::refresh_from(const VelMG<double> &d, hipStream_t s) {
  vgen++;
  mu = d.mu;
  dt = d.dt;
  const int nl = (int)L.size();
  if ((int)d.L.size() != nl || d.nc != nc) {
    set_error("velocity multigrid hierarchies differ");
    return I2E_ERR_ARG;
  }
  auto cast = [&](const DBuf<double> &a, DBuf<T> &b, size_t n) {
    hipLaunchKernelGGL(k_cast<T>, dim3(nblk(n, 256)), dim3(256), 0, s, (int)n, a.p, b.p);
  };
  const int N0 = L[0].N;
  cast(d.L[0].nu, L[0].nu, N0);
  cast(d.L[0].nd, L[0].nd, N0);
  cast(d.L[0].ro, L[0].ro, N0);
  cast(d.L[0].gp, L[0].gp, N0);
  cast(d.L[0].dx, L[0].dx, N0);
  cast(d.L[0].dy, L[0].dy, N0);
  if (diff0)
    hipLaunchKernelGGL(k_sdiff<T>, dim3(nblk((size_t)2 * N0, 256)), dim3(256), 0, s, d.L[0].S.p, N0, L[0].S.p);
  else
    cast(d.L[0].S, L[0].S, (size_t)2 * N0 * 18);
  for (int l = 1; l < nl; l++) cast(d.L[l].S, L[l].S, (size_t)2 * L[l].N * 50);
  // behind d's inversion when that runs on d's second stream (joined with it)
  HIP_OK(hipMemcpyAsync(ainv.p, d.ainv.p, sizeof(double) * (size_t)nc * nc, hipMemcpyDeviceToDevice,
                        d.side ? d.side : s));
  HIP_OK(hipMemcpyAsync(ainvT.p, d.ainvT.p, sizeof(double) * (size_t)nc * nc, hipMemcpyDeviceToDevice,
                        d.side ? d.side : s));
  HIP_OK(hipGetLastError());
  return I2E_OK;
}

// the two coarsest levels in one k_vtail launch: second-coarsest level l
// with the 5x5 Galerkin stencil, small enough for one workgroup; FP32 only
// (the FP64 instance needs every register and spills: 95 us against ~45 us
// for the launches it replaces at 17^2)
template <class T>
bool VelMG<T>::tail_at(int l) const {
  const int nl = (int)L.size();
  return sizeof(T) == 4 && l >= 1 && l == nl - 2 && L[l].W == 2 && L[l].X <= VT_S && L[l].Y <= VT_S && L[l].N <= VT_B &&
         L[nl - 1].N <= VT_C;
}

template <class T>
int VelMG<T>::join(hipStream_t s) {
  if (!side) return I2E_OK;
  HIP_OK(hipEventRecord(ev_join, side));
  HIP_OK(hipStreamWaitEvent(s, ev_join, 0));
  side = nullptr;
  return I2E_OK;
}

template <class T>
void tile_shape(int k, int *sx, int *sy) {
  switch (k) {
    case 1: *sx = ShapeCand<T, 1>::SX, *sy = ShapeCand<T, 1>::SY; break;
    case 2: *sx = ShapeCand<T, 2>::SX, *sy = ShapeCand<T, 2>::SY; break;
    case 3: *sx = ShapeCand<T, 3>::SX, *sy = ShapeCand<T, 3>::SY; break;
    default: *sx = ShapeCand<T, 0>::SX, *sy = ShapeCand<T, 0>::SY; break;
  }
}
template void tile_shape<float>(int, int *, int *);
template void tile_shape<double>(int, int *, int *);

// Galerkin-level smoothing pass with tile shape candidate K
template <class T, bool RS, int K>
static void launch_tw2(const TileArgs<T> &a, const T *S, hipStream_t s, KTimer *kt, int kid, double alg,
                       double rep) {
  constexpr int SX = ShapeCand<T, K>::SX, SY = ShapeCand<T, K>::SY;
  const dim3 TG((a.Y + SX - 1) / SX, (a.X + SY - 1) / SY), TBK(SX * SY);
  KT_LAUNCH(kt, kid, alg, rep, (k_tsmooth<T, T, 2, false, RS, SX, SY>), TG, TBK, 0, s, a, S);
}
template <class T, bool RS>
static void launch_tw2(int k, const TileArgs<T> &a, const T *S, hipStream_t s, KTimer *kt, int kid, double alg,
                       double rep) {
  switch (k) {
    case 1: launch_tw2<T, RS, 1>(a, S, s, kt, kid, alg, rep); break;
    case 2: launch_tw2<T, RS, 2>(a, S, s, kt, kid, alg, rep); break;
    case 3: launch_tw2<T, RS, 3>(a, S, s, kt, kid, alg, rep); break;
    default: launch_tw2<T, RS, 0>(a, S, s, kt, kid, alg, rep); break;
  }
}

// tuned shape per (precision, level size, sweeps), shared by every context
// of the process (an ensemble tunes each level size once)
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int>, int> g_tuned;

template <class T>
int VelMG<T>::tune(hipStream_t s, int reps) {
  const int nl = (int)L.size();
  shape.assign(nl, 0);
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  // levels of the cooperative launch keep its 32x16 tiles
  const int lcoop = coop_level();
  for (int l = 1; l < nl - 1; l++) {
    if (tail_at(l) || (lcoop > 0 && l >= lcoop)) continue;
    VelLevel<T> &v = L[l];
    const int ns = nsmooth_c <= 0 ? nsmooth : nsmooth_c;
    const auto key = std::make_tuple((int)sizeof(T), v.X, v.Y, ns);
    {
      std::lock_guard<std::mutex> g(g_tune_mu);
      auto it = g_tuned.find(key);
      if (it != g_tuned.end()) {
        shape[l] = it->second;
        continue;
      }
    }
    // the pre-smoothing pass from the level's current iterate into its
    // second buffers (nothing the V-cycle reads is written)
    const MapView none{nullptr, nullptr, nullptr, nullptr};
    TileArgs<T> A{v.X, v.Y, v.kx.p, v.ky.p, v.fx.p, v.fy.p, v.ux.p, v.uy.p, v.ux2.p, v.uy2.p, nullptr, nullptr, 0,
                  none, none, none, none, 0, ns, 0, v.dx.p, v.dy.p};
    A.omega = omega;
    double best = 1e300;
    int bk = 0;
    HIP_OK(hipGetLastError());
    for (int k = 0; k < kShapes; k++) {
      for (int w = 0; w < 2; w++) launch_tw2<T, false>(k, A, v.S.p, s, nullptr, 0, 0, 0);
      HIP_OK(hipEventRecord(e0, s));
      for (int r = 0; r < reps; r++) launch_tw2<T, false>(k, A, v.S.p, s, nullptr, 0, 0, 0);
      HIP_OK(hipEventRecord(e1, s));
      HIP_OK(hipEventSynchronize(e1));
      // a shape the device cannot launch (registers, LDS) is no candidate
      const hipError_t le = hipGetLastError();
      if (getenv("I2E_TUNE_VERBOSE"))
        fprintf(stderr, "tune: %s level %dx%d shape %d: %s\n", sizeof(T) == 4 ? "fp32" : "fp64", v.X, v.Y, k,
                hipGetErrorString(le));
      if (le != hipSuccess) continue;
      float ms = 0;
      HIP_OK(hipEventElapsedTime(&ms, e0, e1));
      if (getenv("I2E_TUNE_VERBOSE")) fprintf(stderr, "tune:   %.4f ms per pass\n", ms / reps);
      if (ms < best) {
        best = ms;
        bk = k;
      }
    }
    shape[l] = bk;
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tuned[key] = bk;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return I2E_OK;
}

template <class T>
int VelMG<T>::vcycle(int l, hipStream_t s) {
  const int nl = (int)L.size();
  VelLevel<T> &v = L[l];
  if (l == nl - 1) {
    hipLaunchKernelGGL(k_dense_solve<T>, dim3(nblk(nc, 4)), dim3(256), 0, s, nc, v.N, ainv.p, v.fx.p, v.fy.p, v.ux.p,
                       v.uy.p);
    return I2E_OK;
  }
  const int ns = (l == 0 || nsmooth_c <= 0) ? nsmooth : nsmooth_c;
  if (l > 0 && l == coop_level()) return vcoarse(l, s);
  if (tail_at(l)) {
    const VelLevel<T> &c = L[nl - 1];
    const VTail<T> a{v.X,      v.Y,      c.X,        c.Y,        ns,         v.kx.p,     v.ky.p,   c.kx.p,
                     c.ky.p,   v.S.p,    v.fx.p,     v.fy.p,     view(v.mxn), view(v.mxm), view(v.myn),
                     view(v.mym), ainvT.p, v.ux.p,    v.uy.p,     omega};
    hipLaunchKernelGGL(k_vtail<T>, dim3(1), dim3(VT_B), 0, s, a);
    return I2E_OK;
  }
  TILE_CONSTS(T);
  const dim3 TG((v.Y + TSX - 1) / TSX, (v.X + TSY - 1) / TSY), TBK(TB);
  const int kid = l == 0 ? K_VSMOOTH0 : K_VSMOOTHC;
  // Restriction inside the next level's pre-smoothing pass: in FP64 it adds
  // 3.8 us to that pass against ~7 us for the separate launch it replaces
  // (14.07 -> 14.28 steps/s at 501^2); in FP32 (32x16 tiles) it adds 6.4 us
  // and gains nothing (15.32 -> 15.12)
  constexpr bool fuse_restrict = sizeof(T) == 8;
  const VLev<T> o = lev(l);
  // algorithmic bytes of one smoothing pass: rhs, input and output iterate
  // (two components each), the row kinds, and the coefficients once per
  // point: Galerkin levels 2 x 50 stencil slots; finest level 2 x 18 slots
  // (FP64 ones for FP32 iterates; in difference form plus the two
  // diagonals), or for the matrix-free rows nu, nd, ro, gp and the diagonals
  double coef;
  if (l > 0)
    coef = 100.0 * sizeof(T);
  else if (!stencil0)
    coef = 6.0 * sizeof(T);
  else if (diff0)
    coef = 38.0 * sizeof(T);
  else if (S0d)
    coef = 36.0 * sizeof(double);
  else
    coef = 36.0 * sizeof(T);
  const double bytes = (6.0 * sizeof(T) + coef) * v.N + 2.0 * v.N;
  // a pre-smoothing pass that restricts its right-hand side also reads the
  // finer level's two residual components once (it writes the rhs instead
  // of reading it)
  const double rbytes = l > 0 ? 2.0 * sizeof(T) * L[l - 1].N : 0.0;
  // algorithmic bytes of a smoothing pass on matrix-free rows: ~40 B per
  // point and sweep in FP32 (iterate and right-hand side of both components,
  // the viscosities and density it needs), 80 in FP64
  const double alg = 10.0 * sizeof(T) * (double)v.N * ns;
  const MapView none{nullptr, nullptr, nullptr, nullptr};
  TileArgs<T> A{v.X, v.Y, v.kx.p, v.ky.p, v.fx.p, v.fy.p, v.ux.p, v.uy.p, v.ux2.p, v.uy2.p, nullptr, nullptr, 0,
                none, none, none, none, 1, ns, 0, v.dx.p, v.dy.p};
  A.omega = omega;
  if (l > 0 && fuse_restrict) {
    // this level's right-hand side: restricted from the finer residual
    // inside the pre-smoothing pass (no separate restriction launch)
    const VelLevel<T> &f = L[l - 1];
    A.rfx = f.rx.p;
    A.rfy = f.ry.p;
    A.fxo = v.fx.p;
    A.fyo = v.fy.p;
    A.Yf = f.Y;
    A.fxn = view(f.mxn);
    A.fxm = view(f.mxm);
    A.fyn = view(f.myn);
    A.fym = view(f.mym);
  }
  auto launch = [&](const TileArgs<T> &a) {
    const double rep = bytes + (a.rfx ? rbytes : 0.0);
    if (l == 0 && !stencil0)
      KT_LAUNCH(kt, kid, alg, rep, k_tsmooth0<T>, TG, TBK, 0, s, a, o, (const T *)v.dx.p, (const T *)v.dy.p);
    else if (l == 0 && diff0)
      KT_LAUNCH(kt, kid, alg, rep, (k_tsmooth<T, T, 1, true>), TG, TBK, 0, s, a, (const T *)v.S.p);
    else if (l == 0 && S0d)
      KT_LAUNCH(kt, kid, alg, rep, (k_tsmooth<T, double, 1>), TG, TBK, 0, s, a, S0d);
    else if (l == 0)
      KT_LAUNCH(kt, kid, alg, rep, (k_tsmooth<T, T, 1>), TG, TBK, 0, s, a, (const T *)v.S.p);
    else if (a.rfx)
      launch_tw2<T, true>(l < (int)shape.size() ? shape[l] : 0, a, (const T *)v.S.p, s, kt, kid, alg, rep);
    else
      launch_tw2<T, false>(l < (int)shape.size() ? shape[l] : 0, a, (const T *)v.S.p, s, kt, kid, alg, rep);
    std::swap(v.ux.p, v.ux2.p);
    std::swap(v.uy.p, v.uy2.p);
  };
  launch(A);
  // residual of the smoothed iterate, restricted to the next level
  TileArgs<T> R = A;
  R.ux = v.ux.p;
  R.uy = v.uy.p;
  R.zero = 0;
  R.rfx = R.rfy = nullptr;
  if (l == 0 && !stencil0)
    hipLaunchKernelGGL(k_vresid<T>, grd2(v.X, v.Y), blk2(), 0, s, o, v.ux.p, v.uy.p, v.fx.p, v.fy.p, v.rx.p, v.ry.p);
  else if (l == 0 && diff0)
    hipLaunchKernelGGL((k_tresid<T, T, 1, true>), TG, TBK, 0, s, R, (const T *)v.S.p, v.rx.p, v.ry.p);
  else if (l == 0 && S0d)
    hipLaunchKernelGGL((k_tresid<T, double, 1>), TG, TBK, 0, s, R, S0d, v.rx.p, v.ry.p);
  else if (l == 0)
    hipLaunchKernelGGL((k_tresid<T, T, 1>), TG, TBK, 0, s, R, (const T *)v.S.p, v.rx.p, v.ry.p);
  else
    hipLaunchKernelGGL((k_tresid<T, T, 2>), TG, TBK, 0, s, R, (const T *)v.S.p, v.rx.p, v.ry.p);
  VelLevel<T> &c = L[l + 1];
  // the coarsest level (and the two-level tail) has no smoothing pass to
  // restrict into
  if (l + 1 == nl - 1 || !fuse_restrict || tail_at(l + 1))
    hipLaunchKernelGGL(k_vrestrict<T>, grd2(c.X, c.Y), blk2(), 0, s, c.X, c.Y, v.Y, view(v.mxn), view(v.mxm),
                       view(v.myn), view(v.mym), c.kx.p, c.ky.p, v.rx.p, v.ry.p, c.fx.p, c.fy.p);
  ST_OK(vcycle(l + 1, s));
  TileArgs<T> P = R;
  P.ux = v.ux.p;
  P.uy = v.uy.p;
  P.ox = v.ux2.p;
  P.oy = v.uy2.p;
  P.ex = c.ux.p;
  P.ey = c.uy.p;
  P.Yc = c.Y;
  P.xn = view(v.mxn);
  P.xm = view(v.mxm);
  P.yn = view(v.myn);
  P.ym = view(v.mym);
  P.rev = 1;
  launch(P);
  return I2E_OK;
}


// ---------------------------------------------------------------- cooperative coarse V-cycle (host)
template <class T>
int VelMG<T>::coop_level() const {
  if (sizeof(T) != 4 || !coop) return -1;
  const int nl = (int)L.size();
  int lt = -1;
  for (int l = 1; l < nl; l++)
    if (tail_at(l)) lt = l;
  if (lt < 0) return -1;
  // blocks the device holds at once for this kernel (one per CU at 512
  // threads and its registers; queried once)
  static int resident = -1;
  if (resident < 0) {
    int per = 0, dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_vcoarse<float>, VC_B, 0) != hipSuccess) {
      (void)hipGetLastError();
      resident = 0;
    } else {
      resident = per * cus;
    }
  }
  auto tiles = [&](int l) { return ((L[l].Y + 31) / 32) * ((L[l].X + 15) / 16); };
  int lc = -1, blocks = 1;
  for (int l = lt - 1; l >= 1 && lt - l <= VC_MAXL; l--) {
    if (L[l].W != 2 || tiles(l) > 32) break;
    if (blocks + tiles(l) > resident) break;
    blocks += tiles(l);
    lc = l;
  }
  return lc;
}

template <class T>
int VelMG<T>::vcoarse(int lc, hipStream_t s) {
  if constexpr (sizeof(T) != 4) {
    return I2E_ERR_ARG;
  } else {
    const int nl = (int)L.size();
    int lt = -1;
    for (int l = 1; l < nl; l++)
      if (tail_at(l)) lt = l;
    VCoarse<T> V{};
    V.nl = lt - lc;
    const MapView none{nullptr, nullptr, nullptr, nullptr};
    int b = 0;
    for (int k = 0; k < V.nl; k++) {
      const int l = lc + k;
      VelLevel<T> &v = L[l];
      const int ns = nsmooth_c <= 0 ? nsmooth : nsmooth_c;
      V.b0[k] = b;
      V.ntj[k] = (v.Y + 31) / 32;
      b += V.ntj[k] * ((v.X + 15) / 16);
      TileArgs<T> A{v.X, v.Y, v.kx.p, v.ky.p, v.fx.p, v.fy.p, v.ux.p, v.uy.p, v.ux2.p, v.uy2.p, nullptr, nullptr,
                    0, none, none, none, none, 1, ns, 0, v.dx.p, v.dy.p};
      A.omega = omega;
      if (k > 0) {
        const VelLevel<T> &f = L[l - 1];
        A.rfx = f.rx.p;
        A.rfy = f.ry.p;
        A.fxo = v.fx.p;
        A.fyo = v.fy.p;
        A.Yf = f.Y;
        A.fxn = view(f.mxn);
        A.fxm = view(f.mxm);
        A.fyn = view(f.myn);
        A.fym = view(f.mym);
      }
      V.pre[k] = A;
      TileArgs<T> Rg = A;
      Rg.ux = v.ux2.p;
      Rg.uy = v.uy2.p;
      Rg.zero = 0;
      V.res[k] = Rg;
      const VelLevel<T> &c = L[l + 1];
      TileArgs<T> P = A;
      P.rfx = P.rfy = nullptr;
      P.fxo = P.fyo = nullptr;
      P.zero = 0;
      P.ux = v.ux2.p;
      P.uy = v.uy2.p;
      P.ox = v.ux.p;
      P.oy = v.uy.p;
      P.ex = c.ux.p;
      P.ey = c.uy.p;
      P.Yc = c.Y;
      P.xn = view(v.mxn);
      P.xm = view(v.mxm);
      P.yn = view(v.myn);
      P.ym = view(v.mym);
      P.rev = 1;
      V.post[k] = P;
      V.S[k] = v.S.p;
      V.rx[k] = v.rx.p;
      V.ry[k] = v.ry.p;
    }
    V.b0[V.nl] = b;
    VelLevel<T> &t = L[lt];
    const VelLevel<T> &c = L[nl - 1];
    const VelLevel<T> &f = L[lt - 1];
    const int nst = nsmooth_c <= 0 ? nsmooth : nsmooth_c;
    V.tail = VTail<T>{t.X,    t.Y,    c.X,    c.Y,    nst,    t.kx.p,      t.ky.p,      c.kx.p,      c.ky.p,
                      t.S.p,  t.fx.p, t.fy.p, view(t.mxn), view(t.mxm), view(t.myn), view(t.mym), ainvT.p,
                      t.ux.p, t.uy.p, omega};
    V.tfx = t.fx.p;
    V.tfy = t.fy.p;
    V.Yf = f.Y;
    V.xn = view(f.mxn);
    V.xm = view(f.mxm);
    V.yn = view(f.myn);
    V.ym = view(f.mym);
    // the arguments change only with the hierarchy (set-up and refresh may
    // reallocate level buffers) and the smoothing options: copied to the
    // device when one of those changed since the last launch, through a
    // pinned host buffer whose previous copy has completed
    const int nsx = nsmooth_c <= 0 ? nsmooth : nsmooth_c;
    if (!vch || vk_gen != vgen || vk_l != lc || vk_ns != nsx || vk_om != omega) {
      if (!vch) {
        ST_OK(vcd.alloc(sizeof(V)));
        HIP_OK(hipHostMalloc((void **)&vch, sizeof(V), hipHostMallocDefault));
        HIP_OK(hipEventCreateWithFlags(&vce, hipEventDisableTiming));
      } else {
        HIP_OK(hipEventSynchronize(vce));
      }
      memcpy(vch, &V, sizeof(V));
      HIP_OK(hipMemcpyAsync(vcd.p, vch, sizeof(V), hipMemcpyHostToDevice, s));
      HIP_OK(hipEventRecord(vce, s));
      vk_gen = vgen;
      vk_l = lc;
      vk_ns = nsx;
      vk_om = omega;
    }
    const VCoarse<T> *Vd = (const VCoarse<T> *)vcd.p;
    void *args[] = {&Vd};
    if (kt) kt->begin(K_VCOARSE, s);
    HIP_OK(hipLaunchCooperativeKernel((const void *)k_vcoarse<T>, dim3(b + 1), dim3(VC_B), args, 0, s));
    if (kt) {
      // algorithmic bytes: each level's pre- and post-smoothing sweeps (10
      // sizeof(T) per point and sweep) and residual; representation: the
      // stencils once, the iterates, right-hand sides and residuals
      double alg = 0, rep = 0;
      for (int l = lc; l <= lt; l++) {
        alg += 2.0 * 10.0 * sizeof(T) * L[l].N * nst;
        rep += (100.0 * sizeof(T) + 10.0 * sizeof(T)) * L[l].N;
      }
      kt->end(K_VCOARSE, s, alg, rep);
    }
    return I2E_OK;
  }
}

template <class T>
void VelMG<T>::release() {
  for (auto &v : L) {
    v.nu.release(); v.nd.release(); v.ro.release(); v.gp.release();
    v.kx.release(); v.ky.release(); v.nbx.release(); v.nby.release(); v.kkx.release(); v.kky.release();
    v.dx.release(); v.dy.release(); v.ux.release(); v.uy.release(); v.fx.release(); v.fy.release();
    v.rx.release(); v.ry.release(); v.ux2.release(); v.uy2.release(); v.S.release();
  }
  overflow.release();
  for (auto &b : gfi) b.release();
  for (auto &b : gfj) b.release();
  dense.release();
  ainv.release();
  ainvT.release();
  vcd.release();
  if (vch) (void)hipHostFree(vch);
  if (vce) (void)hipEventDestroy(vce);
  vch = nullptr;
  vce = nullptr;
  vk_gen = -1;
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  ev_fork = ev_join = nullptr;
  side = nullptr;
}

// ---------------------------------------------------------------- scalar MG
template <class T>
POp<T> ScalMG<T>::op(int l) const {
  const ScalLevel<T> &v = L[l];
  return POp<T>{v.X, v.Y, v.cx.p, v.cy.p, v.k.p, v.c.p, v.mask.p, v.dm.p};
}

template <class T>
struct FSplit {
  const T *u;
  int Y;
  __device__ __forceinline__ T operator()(int i, int j) const { return u[i * Y + j]; }
};
template <class T>
struct FUnit {
  int idx, Y;
  __device__ __forceinline__ T operator()(int i, int j) const { return (i * Y + j == idx) ? T(1) : T(0); }
};

template <class T>
__global__ void k_presid(POp<T> o, const T *u, const T *f, T *r) {
  IDX2;
  if (i >= o.X || j >= o.Y) return;
  const int m = i * o.Y + j;
  r[m] = o.mask[m] ? T(0) : f[m] - prow(o, FSplit<T>{u, o.Y}, i, j);
}
template <class T>
__global__ void k_prestrict(int Xc, int Yc, int Yf, MapView mx, MapView my, const int8_t *cmask, const T *r, T *f) {
  IDX2;
  if (i >= Xc || j >= Yc) return;
  double s = 0, ws = 0;
  if (!cmask[i * Yc + j])
    for (int k = mx.rb[i]; k < mx.re[i]; k++) {
      double wx = wgt(mx, k, i);
      if (wx == 0) continue;
      for (int l = my.rb[j]; l < my.re[j]; l++) {
        double w = wx * wgt(my, l, j);
        if (w == 0) continue;
        s += w * (double)r[k * Yf + l];
        ws += w;
      }
    }
  f[i * Yc + j] = T(ws > 0 ? s / ws : 0.0);
}
template <class T>
__global__ void k_pdense(POp<T> o, int n, double *aug) {
  int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (c >= n) return;
  aug[(size_t)r * 2 * n + c] = (double)prow(o, FUnit<T>{c, o.Y}, r / o.Y, r % o.Y);
  aug[(size_t)r * 2 * n + n + c] = (r == c) ? 1.0 : 0.0;
}
// coarse mask by injection, coarse dm = 1 (homogeneous rows)
template <class T>
__global__ void k_pmask_inject(int Xc, int Yc, int Yf, const int *fi, const int *fj, const int8_t *fm, int8_t *cm,
                               T *cdm) {
  IDX2;
  if (i >= Xc || j >= Yc) return;
  cm[i * Yc + j] = fm[fi[i] * Yf + fj[j]];
  cdm[i * Yc + j] = T(1);
}
__global__ void k_copy_mask(int n, const int8_t *a, int8_t *b) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

template <class T>
int ScalMG<T>::setup(const std::vector<double> &x0, const std::vector<double> &y0, int coarse_max) {
  L.clear();
  std::vector<std::vector<double>> xs{x0}, ys{y0};
  std::vector<std::vector<int>> fis, fjs;
  while (true) {
    int X = (int)xs.back().size(), Y = (int)ys.back().size();
    if (std::min(X, Y) - 1 <= coarse_max || std::min(X, Y) < 9) break;
    std::vector<int> ci = coarse_lines(X), cj = coarse_lines(Y);
    std::vector<double> cx, cy;
    for (int a : ci) cx.push_back(xs.back()[a]);
    for (int a : cj) cy.push_back(ys.back()[a]);
    fis.push_back(ci);
    fjs.push_back(cj);
    xs.push_back(cx);
    ys.push_back(cy);
  }
  const int nl = (int)xs.size();
  L.resize(nl);
  sfi.resize(nl);
  sfj.resize(nl);
  for (int l = 0; l < nl; l++) {
    ScalLevel<T> &v = L[l];
    v.X = (int)xs[l].size();
    v.Y = (int)ys[l].size();
    v.N = v.X * v.Y;
    v.hcx = xs[l];
    v.hcy = ys[l];
    ST_OK(v.cx.alloc(v.X));
    ST_OK(v.cy.alloc(v.Y));
    HIP_OK(hipMemcpy(v.cx.p, v.hcx.data(), v.X * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(v.cy.p, v.hcy.data(), v.Y * sizeof(double), hipMemcpyHostToDevice));
    ST_OK(v.k.alloc(v.N));
    ST_OK(v.c.alloc(v.N));
    ST_OK(v.dm.alloc(v.N));
    ST_OK(v.mask.alloc(v.N));
    ST_OK(v.u.alloc(v.N));
    ST_OK(v.f.alloc(v.N));
    ST_OK(v.r.alloc(v.N));
    ST_OK(v.u2.alloc(v.N));
    if (l + 1 < nl) {
      ST_OK(v.mx.build(xs[l], xs[l + 1]));
      ST_OK(v.my.build(ys[l], ys[l + 1]));
      ST_OK(sfi[l].alloc(fis[l].size()));
      ST_OK(sfj[l].alloc(fjs[l].size()));
      HIP_OK(hipMemcpy(sfi[l].p, fis[l].data(), fis[l].size() * sizeof(int), hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(sfj[l].p, fjs[l].data(), fjs[l].size() * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  nc = L[nl - 1].N;
  ST_OK(dense.alloc((size_t)nc * 2 * nc));
  ST_OK(ainv.alloc((size_t)nc * nc));
  return I2E_OK;
}

template <class T>
int ScalMG<T>::refresh(const double *k, const double *c, const int8_t *mask, const double *dm, hipStream_t s) {
  const int nl = (int)L.size();
  const int N0 = L[0].N;
  hipLaunchKernelGGL(k_cast<T>, dim3(nblk(N0, 256)), dim3(256), 0, s, N0, k, L[0].k.p);
  hipLaunchKernelGGL(k_cast<T>, dim3(nblk(N0, 256)), dim3(256), 0, s, N0, c, L[0].c.p);
  hipLaunchKernelGGL(k_cast<T>, dim3(nblk(N0, 256)), dim3(256), 0, s, N0, dm, L[0].dm.p);
  hipLaunchKernelGGL(k_copy_mask, dim3(nblk(N0, 256)), dim3(256), 0, s, N0, mask, L[0].mask.p);
  for (int l = 0; l + 1 < nl; l++) {
    ScalLevel<T> &f = L[l], &cc = L[l + 1];
    hipLaunchKernelGGL(k_restrict_node<T>, grd2(cc.X, cc.Y), blk2(), 0, s, cc.X, cc.Y, f.Y, view(f.mx), view(f.my),
                       f.k.p, cc.k.p, 0);
    hipLaunchKernelGGL(k_restrict_node<T>, grd2(cc.X, cc.Y), blk2(), 0, s, cc.X, cc.Y, f.Y, view(f.mx), view(f.my),
                       f.c.p, cc.c.p, 0);
    hipLaunchKernelGGL(k_pmask_inject<T>, grd2(cc.X, cc.Y), blk2(), 0, s, cc.X, cc.Y, f.Y, sfi[l].p, sfj[l].p,
                       f.mask.p, cc.mask.p, cc.dm.p);
  }
  hipLaunchKernelGGL(k_pdense<T>, dim3(nblk(nc, 256), nc), dim3(256), 0, s, op(nl - 1), nc, dense.p);
  gauss_jordan(nc, dense.p, ainv.p, s);
  HIP_OK(hipGetLastError());
  return I2E_OK;
}

template <class T>
struct PTArgs {
  POp<T> o;
  const T *f;
  const T *u;  // input iterate (not read when zero != 0)
  T *out;      // output iterate
  const T *e;  // correction on the next coarser level, added on load (nullptr: none)
  int Yc;
  MapView mx, my;
  int zero, nsmooth, rev;
  double omega;  // damping of the colour steps (pointwise rows undamped)
};

// Tiled red-black Gauss-Seidel on the scalar 5-point rows (the velocity
// smoother's tile scheme: one launch per pass, halo frozen at its loaded
// value, correction added on load).  Each thread forms its row once in
// difference form, r = c u + sum_n a_n (u - u_n) with a_n = k_face / (h_n h_avg)
// (prow's arithmetic regrouped), pointwise rows r = dm u.  Lanes outside the
// grid compute with zero coefficients and skip the store.
template <class T>
__global__ __launch_bounds__(TileShape<T>::B) void k_ptsmooth(PTArgs<T> A) {
  TILE_CONSTS(T);
  __shared__ T lu[TWX * TWY];
  const POp<T> &o = A.o;
  const int X = o.X, Y = o.Y;
  const int i0 = blockIdx.y * TSY - TH, j0 = blockIdx.x * TSX - TH;
  for (int q = threadIdx.x; q < TWX * TWY; q += TB) {
    const int i = i0 + q / TWX, j = j0 + q % TWX;
    T v = 0;
    if (i >= 0 && i < X && j >= 0 && j < Y) {
      const int m = i * Y + j;
      if (!A.zero) v = A.u[m];
      if (A.e && !o.mask[m]) v += T(prolong_at(A.e, A.Yc, A.mx, A.my, i, j));
    }
    lu[q] = v;
  }
  const int i = i0 + TH + (int)threadIdx.x / TSX, j = j0 + TH + (int)threadIdx.x % TSX;
  const int li = (i - i0) * TWX + (j - j0);
  const bool ok = i < X && j < Y;
  const int m = ok ? i * Y + j : 0;
  T cw = 0, aL = 0, aR = 0, aT = 0, aB = 0, f = 0, rd = 1;
  if (ok) {
    f = A.f[m];
    if (o.mask[m]) {
      cw = o.dm[m];
      rd = T(1) / cw;
    } else {
      const double kc = (double)o.k[m];
      double d = (double)o.c[m];
      if (i > 0 && i < X - 1) {
        const double hL = o.x[i] - o.x[i - 1], hR = o.x[i + 1] - o.x[i], hx = (hL + hR) / 2.0;
        const double l = ((double)o.k[m - Y] + kc) / 2.0 / hL / hx, r = ((double)o.k[m + Y] + kc) / 2.0 / hR / hx;
        aL = T(l);
        aR = T(r);
        d += l + r;
      }
      if (j > 0 && j < Y - 1) {
        const double hT = o.y[j] - o.y[j - 1], hB = o.y[j + 1] - o.y[j], hy = (hT + hB) / 2.0;
        const double t = ((double)o.k[m - 1] + kc) / 2.0 / hT / hy, b = ((double)o.k[m + 1] + kc) / 2.0 / hB / hy;
        aT = T(t);
        aB = T(b);
        d += t + b;
      }
      cw = o.c[m];
      rd = T(A.omega / d);
    }
  }
  __syncthreads();
  for (int it = 0; it < A.nsmooth; it++)
    for (int st = 0; st < 2; st++) {
      const int col = A.rev ? 1 - st : st;
      const bool act = ((i + j) & 1) == col;
      T nv = 0;
      if (act) {
        const T uc = lu[li];
        T r = cw * uc;
        r = fmadd(aL, uc - lu[li - TWX], r);
        r = fmadd(aR, uc - lu[li + TWX], r);
        r = fmadd(aT, uc - lu[li - 1], r);
        r = fmadd(aB, uc - lu[li + 1], r);
        nv = uc + (f - r) * rd;
      }
      __syncthreads();
      if (act && ok) lu[li] = nv;
      __syncthreads();
    }
  if (ok) A.out[m] = lu[li];
}

template <class T>
int ScalMG<T>::vcycle(int l, hipStream_t s) {
  const int nl = (int)L.size();
  ScalLevel<T> &v = L[l];
  if (l == nl - 1) {
    hipLaunchKernelGGL(k_dense_solve<T>, dim3(nblk(nc, 4)), dim3(256), 0, s, nc, nc, ainv.p, v.f.p, v.f.p, v.u.p,
                       v.u.p);
    return I2E_OK;
  }
  const dim3 G = grd2(v.X, v.Y), B = blk2();
  TILE_CONSTS(T);
  const dim3 TG((v.Y + TSX - 1) / TSX, (v.X + TSY - 1) / TSY), TBK(TB);
  const POp<T> o = op(l);
  const MapView none{nullptr, nullptr, nullptr, nullptr, nullptr};
  // pre-smoothing from a zero iterate into u
  hipLaunchKernelGGL(k_ptsmooth<T>, TG, TBK, 0, s,
                     PTArgs<T>{o, v.f.p, nullptr, v.u.p, nullptr, 0, none, none, 1, nsmooth, 0, omega});
  hipLaunchKernelGGL(k_presid<T>, G, B, 0, s, o, v.u.p, v.f.p, v.r.p);
  ScalLevel<T> &c = L[l + 1];
  hipLaunchKernelGGL(k_prestrict<T>, grd2(c.X, c.Y), B, 0, s, c.X, c.Y, v.Y, view(v.mx), view(v.my), c.mask.p, v.r.p,
                     c.f.p);
  ST_OK(vcycle(l + 1, s));
  // coarse-grid correction added on load, post-smoothing into u2, which
  // becomes u (the V-cycle's result stays in L[l].u)
  hipLaunchKernelGGL(k_ptsmooth<T>, TG, TBK, 0, s,
                     PTArgs<T>{o, v.f.p, v.u.p, v.u2.p, c.u.p, c.Y, view(v.mx), view(v.my), 0, nsmooth, 1, omega});
  std::swap(v.u.p, v.u2.p);
  return I2E_OK;
}

template <class T>
void ScalMG<T>::release() {
  for (auto &v : L) {
    v.cx.release(); v.cy.release(); v.k.release(); v.c.release(); v.dm.release();
    v.mask.release(); v.u.release(); v.f.release(); v.r.release(); v.u2.release();
  }
  for (auto &b : sfi) b.release();
  for (auto &b : sfj) b.release();
  dense.release();
  ainv.release();
}

// ---------------------------------------------------------------- cell MG
template <class T>
__global__ void k_cgalerkin(const double *Sf, const int8_t *fmask, int Xf, int Yf, int Xc, int Yc, MapView mx,
                            MapView my, double *Sc, int8_t *cmask, int *overflow) {
  IDX2;
  if (i >= Xc || j >= Yc) return;
  const int Nc = Xc * Yc, Nf = Xf * Yf, m = i * Yc + j;
  double acc[9];
  for (int q = 0; q < 9; q++) acc[q] = 0.0;
  double wsum = 0, weq = 0;
  for (int k = mx.rb[i]; k < mx.re[i]; k++) {
    const double wx = wgt(mx, k, i);
    if (wx == 0) continue;
    for (int l = my.rb[j]; l < my.re[j]; l++) {
      const double w = wx * wgt(my, l, j);
      if (w == 0) continue;
      wsum += w;
      const int a = k * Yf + l;
      if (fmask[a]) continue;
      weq += w;
      for (int dk = -1; dk <= 1; dk++)
        for (int dl = -1; dl <= 1; dl++) {
          const double sv = Sf[(size_t)((dk + 1) * 3 + dl + 1) * Nf + a];
          if (sv == 0) continue;
          const int kb = k + dk, lb = l + dl;
          if (kb < 0 || kb >= Xf || lb < 0 || lb >= Yf) continue;
          const int a0 = mx.i0[kb], b0 = my.i0[lb];
          const double wa = mx.w[kb], wb = my.w[lb];
          for (int ia = 0; ia < 2; ia++)
            for (int ib = 0; ib < 2; ib++) {
              const double pw = (ia ? wa : 1.0 - wa) * (ib ? wb : 1.0 - wb);
              if (pw == 0) continue;
              const int di = a0 + ia - i, dj = b0 + ib - j;
              if (di < -1 || di > 1 || dj < -1 || dj > 1) {
                atomicAdd(overflow, 1);
                continue;
              }
              acc[(di + 1) * 3 + dj + 1] += w * sv * pw;
            }
        }
    }
  }
  const bool id = !(weq > 0);
  cmask[m] = id ? 1 : 0;
  const double inv = wsum > 0 ? 1.0 / wsum : 0.0;
  for (int q = 0; q < 9; q++) Sc[(size_t)q * Nc + m] = id ? (q == 4 ? 1.0 : 0.0) : acc[q] * inv;
}

__global__ void k_cdense(const double *S, int X, int Y, double *aug) {
  const int n = X * Y;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double *row = aug + (size_t)r * 2 * n;
  for (int c = 0; c < 2 * n; c++) row[c] = 0.0;
  row[n + r] = 1.0;
  const int i = r / Y, j = r % Y;
  for (int di = -1; di <= 1; di++)
    for (int dj = -1; dj <= 1; dj++) {
      const int a = i + di, b = j + dj;
      if (a < 0 || a >= X || b < 0 || b >= Y) continue;
      row[a * Y + b] += S[(size_t)((di + 1) * 3 + dj + 1) * n + r];
    }
}

template <class T>
struct CellArgs {
  int X, Y;
  const T *S;
  const int8_t *mask;
  const T *f;
  const T *u;  // input iterate (not read when zero)
  T *o;        // output iterate
  const T *e;  // correction on the next coarser level (nullptr: none)
  int Yc;
  MapView mx, my;
  int zero, nsmooth, rev;
};

// tiled colour-Jacobi smoothing of a 3x3-stencil level (see the velocity
// smoother for the tile scheme); the correction from the next coarser level
// is added on load on the way up
template <class T>
__global__ __launch_bounds__(TileShape<T>::B) void k_ctsmooth(CellArgs<T> A) {
  TILE_CONSTS(T);
  __shared__ T lu[TWX * TWY];
  const int i0 = blockIdx.y * TSY - TH, j0 = blockIdx.x * TSX - TH;
  const int X = A.X, Y = A.Y, N = X * Y;
  for (int q = threadIdx.x; q < TWX * TWY; q += TB) {
    const int i = i0 + q / TWX, j = j0 + q % TWX;
    T v = 0;
    if (i >= 0 && i < X && j >= 0 && j < Y) {
      const int m = i * Y + j;
      if (!A.zero) v = A.u[m];
      if (A.e && !A.mask[m]) v += T(prolong_at(A.e, A.Yc, A.mx, A.my, i, j));
    }
    lu[q] = v;
  }
  const int i = i0 + TH + (int)threadIdx.x / TSX, j = j0 + TH + (int)threadIdx.x % TSX;
  const int la = i - i0, lb = j - j0, li = la * TWX + lb;
  const bool ok = i < X && j < Y;
  const int m = ok ? i * Y + j : 0;
  T sv[9];
#pragma unroll
  for (int q = 0; q < 9; q++) sv[q] = ok ? A.S[(size_t)q * N + m] : T(0);
  const T f = ok ? A.f[m] : T(0), rd = T(1) / (ok ? sv[4] : T(1));
  __syncthreads();
  // out-of-grid lanes compute too and skip the store (dense lane masks; see
  // the velocity smoother)
  for (int it = 0; it < A.nsmooth; it++)
    for (int st = 0; st < 2; st++) {
      const int col = A.rev ? 1 - st : st;
      const bool act = ((i + j) & 1) == col;
      T nv = 0;
      if (act) {
        T acc = 0;
#pragma unroll
        for (int di = -1; di <= 1; di++)
#pragma unroll
          for (int dj = -1; dj <= 1; dj++) acc = fmadd(sv[(di + 1) * 3 + dj + 1], lu[(la + di) * TWX + lb + dj], acc);
        nv = lu[li] + (f - acc) * rd;
      }
      __syncthreads();
      if (act && ok) lu[li] = nv;
      __syncthreads();
    }
  if (ok) A.o[m] = lu[li];
}

// r = f - S u on equation rows (0 on identity rows)
template <class T>
__global__ void k_cresid(int X, int Y, const T *S, const int8_t *mask, const T *u, const T *f, T *r) {
  IDX2;
  if (i >= X || j >= Y) return;
  const int N = X * Y, m = i * Y + j;
  if (mask[m]) {
    r[m] = 0;
    return;
  }
  T acc = 0;
#pragma unroll
  for (int di = -1; di <= 1; di++)
#pragma unroll
    for (int dj = -1; dj <= 1; dj++) {
      int a = i + di, b = j + dj;
      a = a < 0 ? 0 : (a >= X ? X - 1 : a);
      b = b < 0 ? 0 : (b >= Y ? Y - 1 : b);
      acc = fmadd(S[(size_t)((di + 1) * 3 + dj + 1) * N + m], u[a * Y + b], acc);
    }
  r[m] = f[m] - acc;
}

template <class T>
__global__ void k_crestrict(int Xc, int Yc, int Yf, MapView mx, MapView my, const int8_t *cmask, const T *r, T *f) {
  IDX2;
  if (i >= Xc || j >= Yc) return;
  const int m = i * Yc + j;
  f[m] = cmask[m] ? T(0) : restrict_rw(r, Yf, mx, my, i, j);
}

// Residual and restriction of one cell level in one kernel: a workgroup
// takes a CR_T x CR_T tile of coarse cells, forms the fine residuals of the
// fine range the tile gathers from in LDS (as k_cresid), then restricts from
// there (as k_crestrict, same operation order: the result is identical).
// The host checks that every tile's fine range fits CR_F x CR_F.
constexpr int CR_T = 8, CR_F = 2 * CR_T + 4;

template <class T>
__global__ __launch_bounds__(256) void k_cresid_restrict(int X, int Y, const T *S, const int8_t *mask, const T *u,
                                                         const T *f, int Xc, int Yc, MapView mx, MapView my,
                                                         const int8_t *cmask, T *fc) {
  __shared__ T rl[CR_F * CR_F];
  __shared__ int lim[4];
  const int ci0 = blockIdx.y * CR_T, cj0 = blockIdx.x * CR_T;
  const int ci1 = min(ci0 + CR_T, Xc), cj1 = min(cj0 + CR_T, Yc);
  if (threadIdx.x == 0) {
    int a0 = X, a1 = 0, b0 = Y, b1 = 0;
    for (int i = ci0; i < ci1; i++)
      if (mx.re[i] > mx.rb[i]) {
        a0 = min(a0, mx.rb[i]);
        a1 = max(a1, mx.re[i]);
      }
    for (int j = cj0; j < cj1; j++)
      if (my.re[j] > my.rb[j]) {
        b0 = min(b0, my.rb[j]);
        b1 = max(b1, my.re[j]);
      }
    lim[0] = a0;
    lim[1] = a1 > a0 ? a1 : a0;
    lim[2] = b0;
    lim[3] = b1 > b0 ? b1 : b0;
  }
  __syncthreads();
  const int fi0 = lim[0], fj0 = lim[2], FX = lim[1] - lim[0], FY = lim[3] - lim[2];
  const int N = X * Y;
  for (int q = threadIdx.x; q < FX * FY; q += blockDim.x) {
    const int i = fi0 + q / FY, j = fj0 + q % FY, m = i * Y + j;
    T r = 0;
    if (!mask[m]) {
      T acc = 0;
#pragma unroll
      for (int di = -1; di <= 1; di++)
#pragma unroll
        for (int dj = -1; dj <= 1; dj++) {
          int a = i + di, b = j + dj;
          a = a < 0 ? 0 : (a >= X ? X - 1 : a);
          b = b < 0 ? 0 : (b >= Y ? Y - 1 : b);
          acc = fmadd(S[(size_t)((di + 1) * 3 + dj + 1) * N + m], u[a * Y + b], acc);
        }
      r = f[m] - acc;
    }
    rl[(q / FY) * CR_F + q % FY] = r;
  }
  __syncthreads();
  const int TY = cj1 - cj0, TN = (ci1 - ci0) * TY;
  for (int q = threadIdx.x; q < TN; q += blockDim.x) {
    const int i = ci0 + q / TY, j = cj0 + q % TY, m = i * Yc + j;
    if (cmask[m]) {
      fc[m] = T(0);
      continue;
    }
    const int k0 = mx.rb[i] - fi0, l0 = my.rb[j] - fj0;
    double wx[RW], wy[RW];
#pragma unroll
    for (int a = 0; a < RW; a++) {
      wx[a] = mx.rw[i * RW + a];
      wy[a] = my.rw[j * RW + a];
    }
    double sum = 0;
#pragma unroll
    for (int a = 0; a < RW; a++)
#pragma unroll
      for (int b = 0; b < RW; b++)
        if (wx[a] != 0 && wy[b] != 0) sum += wx[a] * wy[b] * (double)rl[(k0 + a) * CR_F + l0 + b];
    fc[m] = T(sum);
  }
}

// The small levels of the cell multigrid (from the first with N <= CT_N
// down to the coarsest) run as one workgroup, phases separated by workgroup
// barriers.  Replaces ~4 launches per level.  Layout for issue efficiency on
// one CU:
//  - each thread's points are taken in colour-sorted order (all red points,
//    then all black), so in a colour step whole wavefronts are idle instead
//    of every wavefront running with half its lanes;
//  - iterates live in LDS with a zero border, so the 3x3 neighbours are
//    fixed offsets (no clamping); right-hand sides, residuals and row kinds
//    in LDS unpadded;
//  - the first tail level's stencil rows stay in registers; the coarsest
//    inverse, the coarser levels' stencils, every level's row kinds and the
//    grid transfer maps arrive as one packed LDS image (CellMG::timg, filled
//    at refresh), copied by a single round of independent 16-byte loads: as
//    separate copy loops they cost one global round trip each, ~20 us;
//  - point indices are formed once.
constexpr int CT_N = 1156, CT_THR = 1024, CT_L = 5, CT_TOT = 1600, CT_PTOT = 2400, CT_IMG = 112 * 1024,
              CT_K = (CT_N + CT_THR - 1) / CT_THR;

template <class T>
struct CTLev {
  int X, Y, N;
  int off, poff;  // LDS offsets of the unpadded arrays (f, r) and of the padded iterate
  // byte offsets in the image: stencil (levels >= 1), row kinds, and the
  // maps to the next coarser level (x: i0, w, rb, rw; then y)
  int s_b, m_b, xi0, xw, xrb, xrw, yi0, yw, yrb, yrw;
  const T *S;  // stencil in device memory (the first tail level's rows go to registers)
};
template <class T>
struct CTArgs {
  int nl, nc, ptot, n16, ainv_b;
  CTLev<T> L[CT_L];
  const uint4 *img;
  const T *fin;
  T *uout;
  int nsmooth;
};

struct TailSeg {
  size_t dst;
  const void *src;
  size_t bytes;
};

// Layout of the tail from level l0: fills A (when given) and the copy list
// of the image (when given); returns the image size in bytes, or 0 when the
// tail does not fit the kernel's LDS arrays.
template <class T>
static size_t tail_layout(const CellMG<T> &g, int l0, CTArgs<T> *A, std::vector<TailSeg> *segs) {
  const int nl = (int)g.L.size() - l0;
  if (nl > CT_L || g.L[l0].N > CT_N || g.nc > GJ_LDS_N) return 0;
  size_t b = 0;
  auto put = [&](const void *src, size_t bytes) {
    b = (b + 15) / 16 * 16;
    const size_t o = b;
    if (segs && bytes) segs->push_back(TailSeg{o, src, bytes});
    b += bytes;
    return (int)o;
  };
  CTArgs<T> a{};
  a.nl = nl;
  a.nc = g.nc;
  a.ainv_b = put(g.ainv.p, sizeof(double) * (size_t)g.nc * g.nc);
  int off = 0, poff = 0;
  for (int q = 0; q < nl; q++) {
    const CellLevel<T> &w = g.L[l0 + q];
    CTLev<T> &c = a.L[q];
    c.X = w.X;
    c.Y = w.Y;
    c.N = w.N;
    c.off = off;
    c.poff = poff;
    c.S = w.S.p;
    off += w.N;
    poff += (w.X + 2) * (w.Y + 2);
    c.s_b = q > 0 ? put(w.S.p, sizeof(T) * 9 * (size_t)w.N) : 0;
    c.m_b = put(w.mask.p, (size_t)w.N);
    if (q + 1 < nl) {
      const int Xc = g.L[l0 + q + 1].X, Yc = g.L[l0 + q + 1].Y;
      c.xi0 = put(w.mx.i0.p, sizeof(int) * w.X);
      c.xw = put(w.mx.w.p, sizeof(double) * w.X);
      c.xrb = put(w.mx.rb.p, sizeof(int) * Xc);
      c.xrw = put(w.mx.rw.p, sizeof(double) * RW * Xc);
      c.yi0 = put(w.my.i0.p, sizeof(int) * w.Y);
      c.yw = put(w.my.w.p, sizeof(double) * w.Y);
      c.yrb = put(w.my.rb.p, sizeof(int) * Yc);
      c.yrw = put(w.my.rw.p, sizeof(double) * RW * Yc);
    }
  }
  b = (b + 15) / 16 * 16;
  a.ptot = poff;
  a.n16 = (int)(b / 16);
  if (b > (size_t)CT_IMG || off > CT_TOT || poff > CT_PTOT) return 0;
  if (A) *A = a;
  return b;
}

// colour-sorted position q (red points (i+j even) first, then black) -> (i, j)
__device__ __forceinline__ void ct_point(int q, int X, int Y, int &i, int &j) {
  const int nred = (X * Y + 1) / 2;
  const bool red = q < nred;
  const int r = red ? q : q - nred;
  const int c0 = red ? (Y + 1) / 2 : Y / 2;  // points of this colour in an even row
  const int pair = r / Y, rem = r % Y;
  const int odd = rem >= c0 ? 1 : 0;
  i = 2 * pair + odd;
  j = 2 * (odd ? rem - c0 : rem) + (red ? (i & 1) : 1 - (i & 1));
}

// bilinear correction from a padded coarse iterate (prolong_at's weights)
template <class T>
__device__ __forceinline__ double ct_prolong(const T *e, int Ycp, const MapView &mx, const MapView &my, int i, int j) {
  const int a = mx.i0[i] + 1, b = my.i0[j] + 1;
  const double wa = mx.w[i], wb = my.w[j];
  return (1 - wa) * ((1 - wb) * (double)e[a * Ycp + b] + wb * (double)e[a * Ycp + b + 1]) +
         wa * ((1 - wb) * (double)e[(a + 1) * Ycp + b] + wb * (double)e[(a + 1) * Ycp + b + 1]);
}

template <class T>
__global__ __launch_bounds__(CT_THR) void k_ctail(CTArgs<T> A) {
  __shared__ uint4 img[CT_IMG / 16];
  __shared__ T u[CT_PTOT], f[CT_TOT], r[CT_TOT];
  const int t = threadIdx.x;
  const char *ib = (const char *)img;
  // the image: four independent 16-byte loads per thread per round (one
  // round up to 64 KB), stored after all of them have been issued
  for (int base = 0; base < A.n16; base += 4 * CT_THR) {
    uint4 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int q = base + t + k * CT_THR;
      v[k] = A.img[q < A.n16 ? q : A.n16 - 1];  // unconditional (in-bounds) loads stay in registers
    }
    // lanes past the end rewrite the last element with its own value
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int q = base + t + k * CT_THR;
      img[q < A.n16 ? q : A.n16 - 1] = v[k];
    }
  }
  for (int q = t; q < A.ptot; q += CT_THR) u[q] = 0;
  const double *ainv = (const double *)(ib + A.ainv_b);
  auto msk = [&](const CTLev<T> &L) { return (const int8_t *)(ib + L.m_b); };
  auto Ss = [&](const CTLev<T> &L) { return (const T *)(ib + L.s_b); };
  // transfer maps of a level (to the next coarser one), read from LDS
  auto lmx = [&](const CTLev<T> &L) {
    return MapView{(const int *)(ib + L.xi0), (const double *)(ib + L.xw), (const int *)(ib + L.xrb), nullptr,
                   (const double *)(ib + L.xrw)};
  };
  auto lmy = [&](const CTLev<T> &L, int) {
    return MapView{(const int *)(ib + L.yi0), (const double *)(ib + L.yw), (const int *)(ib + L.yrb), nullptr,
                   (const double *)(ib + L.yrw)};
  };
  for (int p = t; p < A.L[0].N; p += CT_THR) f[A.L[0].off + p] = A.fin[p];
  // this thread's points per level: linear index (-1: none) and padded index
  int lp[CT_L][CT_K], pp[CT_L][CT_K];
#pragma unroll
  for (int l = 0; l < CT_L; l++)
#pragma unroll
    for (int k = 0; k < CT_K; k++) {
      lp[l][k] = -1;
      pp[l][k] = 0;
      const int q = t + k * CT_THR;
      if (l < A.nl && q < A.L[l].N) {
        int i, j;
        ct_point(q, A.L[l].X, A.L[l].Y, i, j);
        lp[l][k] = i * A.L[l].Y + j;
        pp[l][k] = A.L[l].poff + (i + 1) * (A.L[l].Y + 2) + j + 1;
      }
    }
  T s0[CT_K][9];
#pragma unroll
  for (int k = 0; k < CT_K; k++)
#pragma unroll
    for (int q = 0; q < 9; q++) s0[k][q] = lp[0][k] >= 0 ? A.L[0].S[(size_t)q * A.L[0].N + lp[0][k]] : T(0);
  __syncthreads();
  // S u at this thread's point k of level l (l compile-time after unrolling)
  auto apply = [&](int l, int k) {
    const CTLev<T> &L = A.L[l];
    const int Yp = L.Y + 2, pc = pp[l][k];
    T acc = 0;
#pragma unroll
    for (int di = -1; di <= 1; di++)
#pragma unroll
      for (int dj = -1; dj <= 1; dj++) {
        const int q = (di + 1) * 3 + dj + 1;
        const T sv = l == 0 ? s0[k][q] : Ss(L)[q * L.N + lp[l][k]];
        acc = fmadd(sv, u[pc + di * Yp + dj], acc);
      }
    return acc;
  };
  auto diag = [&](int l, int k) { return l == 0 ? s0[k][4] : Ss(A.L[l])[4 * A.L[l].N + lp[l][k]]; };
  auto smooth = [&](int l, bool rev) {
    const CTLev<T> &L = A.L[l];
    const int nred = (L.N + 1) / 2;
    for (int it = 0; it < A.nsmooth; it++)
      for (int st = 0; st < 2; st++) {
        const int col = rev ? 1 - st : st;
        T nv[CT_K];
        bool act[CT_K];
#pragma unroll
        for (int k = 0; k < CT_K; k++) {
          const int q = t + k * CT_THR;
          act[k] = q < L.N && (q >= nred ? 1 : 0) == col;
          nv[k] = 0;
          if (act[k]) nv[k] = u[pp[l][k]] + (f[L.off + lp[l][k]] - apply(l, k)) / diag(l, k);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < CT_K; k++)
          if (act[k]) u[pp[l][k]] = nv[k];
        __syncthreads();
      }
  };
  // down: pre-smooth, residual, restriction
#pragma unroll
  for (int l = 0; l < CT_L - 1; l++) {
    if (l >= A.nl - 1) break;
    const CTLev<T> &L = A.L[l], &C = A.L[l + 1];
    smooth(l, false);
#pragma unroll
    for (int k = 0; k < CT_K; k++)
      if (lp[l][k] >= 0) {
        const int m = L.off + lp[l][k];
        r[m] = msk(L)[lp[l][k]] ? T(0) : f[m] - apply(l, k);
      }
    __syncthreads();
    const MapView mx = lmx(L), my = lmy(L, C.X);
    for (int p = t; p < C.N; p += CT_THR) {
      const int i = p / C.Y, j = p % C.Y;
      f[C.off + p] = msk(C)[p] ? T(0) : restrict_rw(r + L.off, L.Y, mx, my, i, j);
    }
    __syncthreads();
  }
  // coarsest: dense FP64 inverse, one wavefront per row
  {
    const CTLev<T> &L = A.L[A.nl - 1];
    const int wv = t >> 6, lane = t & 63, nw = CT_THR / 64;
    for (int row = wv; row < A.nc; row += nw) {
      double acc = 0;
      for (int c = lane; c < A.nc; c += 64) acc += ainv[row * A.nc + c] * (double)f[L.off + c];
      acc = wave_sum(acc);
      if (lane == 0) u[L.poff + (row / L.Y + 1) * (L.Y + 2) + row % L.Y + 1] = T(acc);
    }
    __syncthreads();
  }
  // up: prolongation, post-smooth
#pragma unroll
  for (int l = CT_L - 2; l >= 0; l--) {
    if (l >= A.nl - 1) continue;
    const CTLev<T> &L = A.L[l], &C = A.L[l + 1];
    const MapView mx = lmx(L), my = lmy(L, C.X);
#pragma unroll
    for (int k = 0; k < CT_K; k++)
      if (lp[l][k] >= 0 && !msk(L)[lp[l][k]]) {
        const int i = lp[l][k] / L.Y, j = lp[l][k] % L.Y;
        u[pp[l][k]] += T(ct_prolong(u + C.poff, C.Y + 2, mx, my, i, j));
      }
    __syncthreads();
    smooth(l, true);
  }
#pragma unroll
  for (int k = 0; k < CT_K; k++)
    if (lp[0][k] >= 0) A.uout[lp[0][k]] = u[pp[0][k]];
}

template <class T>
int CellMG<T>::setup(const std::vector<double> &x0, const std::vector<double> &y0, int coarse_max) {
  L.clear();
  std::vector<std::vector<double>> xs{x0}, ys{y0};
  while (true) {
    int X = (int)xs.back().size(), Y = (int)ys.back().size();
    if (std::min(X, Y) - 1 <= coarse_max || std::min(X, Y) < 9) break;
    std::vector<int> ci = coarse_lines(X), cj = coarse_lines(Y);
    std::vector<double> cx, cy;
    for (int a : ci) cx.push_back(xs.back()[a]);
    for (int a : cj) cy.push_back(ys.back()[a]);
    xs.push_back(cx);
    ys.push_back(cy);
  }
  const int nl = (int)xs.size();
  L.resize(nl);
  for (int l = 0; l < nl; l++) {
    CellLevel<T> &v = L[l];
    v.X = (int)xs[l].size();
    v.Y = (int)ys[l].size();
    v.N = v.X * v.Y;
    ST_OK(v.S.alloc((size_t)9 * v.N));
    ST_OK(v.mask.alloc(v.N));
    ST_OK(v.u.alloc(v.N));
    ST_OK(v.u2.alloc(v.N));
    ST_OK(v.f.alloc(v.N));
    ST_OK(v.r.alloc(v.N));
    if (l + 1 < nl) {
      ST_OK(v.mx.build(xs[l], xs[l + 1]));
      ST_OK(v.my.build(ys[l], ys[l + 1]));
      v.fused = v.mx.tile_span(CR_T) <= CR_F && v.my.tile_span(CR_T) <= CR_F;
    }
  }
  nc = L[nl - 1].N;
  ST_OK(dense.alloc((size_t)nc * 2 * nc));
  ST_OK(ainv.alloc((size_t)nc * nc));
  ST_OK(overflow.alloc(1));
  HIP_OK(hipMemset(overflow.p, 0, sizeof(int)));
  // the tail takes over at the first level that fits it
  tail_l = -1;
  for (int l = 0; l < nl; l++) {
    const size_t b = tail_layout(*this, l, (CTArgs<T> *)nullptr, nullptr);
    if (b) {
      tail_l = l;
      ST_OK(timg.alloc(b / 16));
      break;
    }
  }
  return I2E_OK;
}

template <class T>
int CellMG<T>::tail_pack(hipStream_t s) {
  if (tail_l < 0) return I2E_OK;
  std::vector<TailSeg> segs;
  tail_layout(*this, tail_l, (CTArgs<T> *)nullptr, &segs);
  for (const TailSeg &g : segs)
    HIP_OK(hipMemcpyAsync((char *)timg.p + g.dst, g.src, g.bytes, hipMemcpyDeviceToDevice, s));
  return I2E_OK;
}

template <class T>
int CellMG<T>::refresh(const double *S0, const int8_t *mask0, hipStream_t s) {
  const int nl = (int)L.size();
  if (sizeof(T) != sizeof(double)) {
    set_error("cell multigrid: form the hierarchy in FP64 and round it with refresh_from");
    return I2E_ERR_ARG;
  }
  HIP_OK(hipMemcpyAsync(L[0].S.p, S0, sizeof(double) * 9 * (size_t)L[0].N, hipMemcpyDeviceToDevice, s));
  HIP_OK(hipMemcpyAsync(L[0].mask.p, mask0, (size_t)L[0].N, hipMemcpyDeviceToDevice, s));
  for (int l = 0; l + 1 < nl; l++) {
    CellLevel<T> &f = L[l], &c = L[l + 1];
    hipLaunchKernelGGL(k_cgalerkin<T>, grd2(c.X, c.Y), blk2(), 0, s, (const double *)f.S.p, f.mask.p, f.X, f.Y, c.X,
                       c.Y, view(f.mx), view(f.my), (double *)c.S.p, c.mask.p, overflow.p);
  }
  hipLaunchKernelGGL(k_cdense, dim3(nblk(nc, 64)), dim3(64), 0, s, (const double *)L[nl - 1].S.p, L[nl - 1].X,
                     L[nl - 1].Y, dense.p);
  gauss_jordan(nc, dense.p, ainv.p, s);
  HIP_OK(hipGetLastError());
  ST_OK(tail_pack(s));
  int nover = 0;
  HIP_OK(hipMemcpyAsync(&nover, overflow.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (nover) {
    set_error("cell multigrid: %d Galerkin couplings fall outside the 3x3 window", nover);
    return I2E_ERR_UNSUPPORTED;
  }
  return I2E_OK;
}

template <class T>
int CellMG<T>::refresh_from(const CellMG<double> &d, hipStream_t s) {
  const int nl = (int)L.size();
  if ((int)d.L.size() != nl || d.nc != nc) {
    set_error("cell multigrid hierarchies differ");
    return I2E_ERR_ARG;
  }
  for (int l = 0; l < nl; l++) {
    const size_t n9 = (size_t)9 * L[l].N;
    hipLaunchKernelGGL(k_cast<T>, dim3(nblk(n9, 256)), dim3(256), 0, s, (int)n9, d.L[l].S.p, L[l].S.p);
    HIP_OK(hipMemcpyAsync(L[l].mask.p, d.L[l].mask.p, (size_t)L[l].N, hipMemcpyDeviceToDevice, s));
  }
  HIP_OK(hipMemcpyAsync(ainv.p, d.ainv.p, sizeof(double) * (size_t)nc * nc, hipMemcpyDeviceToDevice, s));
  HIP_OK(hipGetLastError());
  return tail_pack(s);
}

template <class T>
int CellMG<T>::vcycle(int l, hipStream_t s) {
  const int nl = (int)L.size();
  CellLevel<T> &v = L[l];
  if (l == tail_l) {
    CTArgs<T> A{};
    if (!tail_layout(*this, l, &A, nullptr)) {
      set_error("cell multigrid tail layout changed");
      return I2E_ERR_ARG;
    }
    A.img = timg.p;
    A.fin = v.f.p;
    A.uout = v.u.p;
    A.nsmooth = (l == 0 || nsmooth_c <= 0) ? nsmooth : nsmooth_c;
    hipLaunchKernelGGL(k_ctail<T>, dim3(1), dim3(CT_THR), 0, s, A);
    return I2E_OK;
  }
  if (l == nl - 1) {
    hipLaunchKernelGGL(k_dense_solve<T>, dim3(nblk(nc, 4)), dim3(256), 0, s, nc, nc, ainv.p, v.f.p, v.f.p, v.u.p,
                       v.u.p);
    return I2E_OK;
  }
  TILE_CONSTS(T);
  const dim3 TG((v.Y + TSX - 1) / TSX, (v.X + TSY - 1) / TSY), TBK(TB);
  const MapView none{nullptr, nullptr, nullptr, nullptr, nullptr};
  const int ns = (l == 0 || nsmooth_c <= 0) ? nsmooth : nsmooth_c;
  CellArgs<T> A{v.X, v.Y, v.S.p, v.mask.p, v.f.p, v.u.p, v.u2.p, nullptr, 0, none, none, 1, ns, 0};
  hipLaunchKernelGGL(k_ctsmooth<T>, TG, TBK, 0, s, A);
  std::swap(v.u.p, v.u2.p);
  CellLevel<T> &c = L[l + 1];
  if (v.fused) {
    hipLaunchKernelGGL(k_cresid_restrict<T>, dim3((c.Y + CR_T - 1) / CR_T, (c.X + CR_T - 1) / CR_T), dim3(256), 0,
                       s, v.X, v.Y, (const T *)v.S.p, (const int8_t *)v.mask.p, (const T *)v.u.p, (const T *)v.f.p,
                       c.X, c.Y, view(v.mx), view(v.my), (const int8_t *)c.mask.p, c.f.p);
  } else {
    hipLaunchKernelGGL(k_cresid<T>, grd2(v.X, v.Y), blk2(), 0, s, v.X, v.Y, (const T *)v.S.p,
                       (const int8_t *)v.mask.p, (const T *)v.u.p, (const T *)v.f.p, v.r.p);
    hipLaunchKernelGGL(k_crestrict<T>, grd2(c.X, c.Y), blk2(), 0, s, c.X, c.Y, v.Y, view(v.mx), view(v.my),
                       (const int8_t *)c.mask.p, (const T *)v.r.p, c.f.p);
  }
  ST_OK(vcycle(l + 1, s));
  CellArgs<T> P{v.X, v.Y, v.S.p, v.mask.p, v.f.p, v.u.p, v.u2.p, c.u.p, c.Y, view(v.mx), view(v.my), 0, ns, 1};
  hipLaunchKernelGGL(k_ctsmooth<T>, TG, TBK, 0, s, P);
  std::swap(v.u.p, v.u2.p);
  return I2E_OK;
}

template <class T>
void CellMG<T>::release() {
  for (auto &v : L) {
    v.S.release(); v.mask.release(); v.u.release(); v.u2.release(); v.f.release(); v.r.release();
  }
  dense.release();
  ainv.release();
  overflow.release();
  timg.release();
}

template struct CellMG<float>;
template struct CellMG<double>;
template struct VelMG<float>;
template struct VelMG<double>;
template struct ScalMG<float>;
template struct ScalMG<double>;

}  // namespace i2e
