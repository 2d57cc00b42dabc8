// band.hip -- the Stokes rows assembled into band storage and factorised by
// LU with partial pivoting inside the band: the FGMRES preconditioner of
// stokes_direct = 1.  The reference solves every Stokes system directly
// (gausmat4 -> PARDISO, gausmart.c:174-404); this is the direct solve a
// banded matrix allows on one GPU, for the systems the multigrid cannot
// precondition (melt at 1e2 Pa s against the 1e23 Pa s clip).
//
// Unknowns keep the reference's order 3 m + {0: P, 1: Vx, 2: Vy}, m = i Y + j
// (headmart.c:87-99), so the rows reach about 3 Y unknowns either side of
// the diagonal.  Steps:
//   1. each row's coefficients, probed from the matrix-free rows (rows.hpp)
//      on unit vectors of the unknowns within two nodes, BC rows from their
//      tables, into a fixed-width row list;
//   2. Ruiz equilibration (ten sweeps of row and column max scaling), since
//      the rows span mukoef = 1e16 times velocities against momentum rows of
//      1e2..1e23 Pa s viscosities;
//   3. column-major band storage with kl sub- and kl + ku superdiagonals (the
//      fill of row interchanges) plus BNB rows for a panel's multipliers:
//      A(i, c) at ab[c LD + KV + i - c], KV = kl + ku, LD = KV + kl + BNB;
//   4. per panel of BNB columns: one workgroup factors the panel with
//      partial pivoting (rows interchanged across the whole panel), applies
//      the interchanges and the unit-lower triangle to the panel rows of the
//      trailing columns; then the trailing band window (kl rows x KV
//      columns) takes the rank-BNB update;
//   5. the solves: one workgroup each, the active part of the vector in an
//      LDS window, panel by panel.
#include <climits>

#include "ctx.hpp"
#include "ops.hpp"
#include "rows.hpp"

namespace i2e {

constexpr int BNE = 40;    // assembled entries per row at most
constexpr int BNB = 16;    // panel width
constexpr int BPT = 1024;  // panel and solve workgroup
constexpr int BROWS = 2;   // panel rows per thread at most: kl + BNB <= BROWS * BPT
constexpr int BW = 4096;   // LDS window of the solves (doubles): kl + BNB, KV + BNB <= BW
constexpr int BSWEEPS = 10;

struct BandLU {
  int64_t n = 0;  // unknowns (3 N)
  int kl = 0, ku = 0, KV = 0;
  int64_t LD = 0;
  DBuf<double> ab, rs, cs, rmx, w, eval;
  DBuf<unsigned long long> cmx;
  DBuf<int> ipiv, ecol, ecnt, dims;
};

void band_release(BandLU *b) {
  if (!b) return;
  b->ab.release();
  b->rs.release();
  b->cs.release();
  b->rmx.release();
  b->w.release();
  b->eval.release();
  b->cmx.release();
  b->ipiv.release();
  b->ecol.release();
  b->ecnt.release();
  b->dims.release();
  delete b;
}

// ---------------------------------------------------------------- assembly
// One thread per node; its three rows exactly as k_stokes_apply evaluates
// them (ops.hip), each coefficient the row's value on a unit vector.
__global__ void k_band_rows(SOp<double> o, BcTab bc, int *ecol, double *eval, int *ecnt, int *dims) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = o.X, Y = o.Y;
  if (j >= Y) return;
  const int m = i * Y + j;
  int lo = 0, hi = 0, anyover = 0;
  for (int comp = 0; comp < 3; comp++) {
    const int u = 3 * m + comp;
    int *rc = ecol + (size_t)u * BNE;
    double *rv = eval + (size_t)u * BNE;
    int n = 0;
    bool over = false;
    auto put = [&](int col, double v, bool merge) {
      if (merge)
        for (int q = 0; q < n; q++)
          if (rc[q] == col) {
            rv[q] += v;
            return;
          }
      if (n >= BNE) {
        over = true;
        return;
      }
      rc[n] = col;
      rv[n] = v;
      n++;
    };
    // f on the unit vector of every unknown within two nodes of (ci, cj)
    auto probe = [&](int ci, int cj, auto f) {
      for (int di = -2; di <= 2; di++)
        for (int dj = -2; dj <= 2; dj++) {
          const int ii = ci + di, jj = cj + dj;
          if (ii < 0 || ii >= X || jj < 0 || jj >= Y) continue;
          for (int cc = 0; cc < 3; cc++) {
            const VUnit<double> e{cc, ii * Y + jj, Y};
            const double v = f(e);
            if (v != 0) put(3 * (ii * Y + jj) + cc, v, false);
          }
        }
    };
    // mukoef (x_u - sum bondv x_nb)
    auto bcrow = [&](int b) {
      put(u, o.mu, true);
      for (int q = 0; q < 3; q++) {
        const int nb = bc.bondn[3 * b + q];
        if (nb) put(nb - 1, -o.mu * bc.bondv[4 * b + q + 1], true);
      }
    };
    auto vbc = [&](int b) {
      if (bc.bondv[4 * b + 1] > 1e30) {
        int dm1, dm2;
        cont_shift(bc.bondn[3 * b] - 1 - u, Y, dm1, dm2);
        const int ci = i + dm1, cj = j + dm2;
        probe(ci, cj, [&](const VUnit<double> &e) { return o.mu * cont(o, e, ci, cj); });
      } else {
        bcrow(b);
      }
    };
    if (comp == 0) {
      if (i && j) {
        if (bc.bondm[u])
          bcrow(bc.bondm[u]);
        else
          probe(i, j, [&](const VUnit<double> &e) { return o.mu * cont(o, e, i, j); });
      } else {
        put(u, o.mu, false);
      }
    } else if (comp == 1) {
      if (j < Y - 1) {
        if (bc.bondm[u])
          vbc(bc.bondm[u]);
        else
          probe(i, j, [&](const VUnit<double> &e) { return xmom(o, e, i, j, true); });
      } else {
        put(u, o.mu, false);
      }
    } else {
      if (i < X - 1) {
        if (bc.bondm[u])
          vbc(bc.bondm[u]);
        else
          probe(i, j, [&](const VUnit<double> &e) { return ymom(o, e, i, j, true); });
      } else {
        put(u, o.mu, false);
      }
    }
    ecnt[u] = n;
    for (int q = 0; q < n; q++) {
      const int d = u - rc[q];
      lo = d > lo ? d : lo;
      hi = -d > hi ? -d : hi;
    }
    anyover |= over;
  }
  atomicMax(&dims[0], lo);
  atomicMax(&dims[1], hi);
  if (anyover) atomicMax(&dims[2], 1);
}

__global__ void k_band_fill(int64_t n, double v, double *a) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    a[k] = v;
}
__global__ void k_band_zero_u64(int64_t n, unsigned long long *a) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    a[k] = 0;
}

// Ruiz sweep: row maxima, column maxima (the bit patterns of non-negative
// doubles order like the values), then B <- diag(1/sqrt(r)) B diag(1/sqrt(c))
__global__ void k_band_rowmax(int64_t n, const int *ecnt, const double *eval, double *rmx) {
  const int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (u >= n) return;
  double mx = 0;
  for (int q = 0; q < ecnt[u]; q++) mx = fmax(mx, fabs(eval[(size_t)u * BNE + q]));
  rmx[u] = mx;
}
__global__ void k_band_colmax(int64_t n, const int *ecnt, const int *ecol, const double *eval,
                              unsigned long long *cmx) {
  const int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (u >= n) return;
  for (int q = 0; q < ecnt[u]; q++)
    atomicMax(&cmx[ecol[(size_t)u * BNE + q]],
              (unsigned long long)__double_as_longlong(fabs(eval[(size_t)u * BNE + q])));
}
__device__ __forceinline__ double ruiz_div(double mx) { return mx > 0 ? sqrt(mx) : 1.0; }
__global__ void k_band_scale(int64_t n, const int *ecnt, const int *ecol, double *eval, const double *rmx,
                             const unsigned long long *cmx, double *rs, double *cs) {
  const int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (u >= n) return;
  const double dr = ruiz_div(rmx[u]);
  for (int q = 0; q < ecnt[u]; q++) {
    const size_t e = (size_t)u * BNE + q;
    eval[e] = eval[e] / dr / ruiz_div(__longlong_as_double((long long)cmx[ecol[e]]));
  }
  rs[u] /= dr;
  cs[u] /= ruiz_div(__longlong_as_double((long long)cmx[u]));
}

__global__ void k_band_scatter(int64_t n, const int *ecnt, const int *ecol, const double *eval, double *ab,
                               int64_t LD, int KV) {
  const int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (u >= n) return;
  for (int q = 0; q < ecnt[u]; q++) {
    const int64_t c = ecol[(size_t)u * BNE + q];
    ab[c * LD + KV + u - c] = eval[(size_t)u * BNE + q];
  }
}

// ---------------------------------------------------------------- factorisation
// Panel columns j0 .. j0+jb-1, rows j0 .. j0+jb+kl-1 (below that the panel
// columns are zero): thread t holds panel rows t and t + BPT in registers.
// Partial pivoting picks the largest magnitude in the column (the first
// such row on ties), rows are interchanged across the whole panel, the
// multipliers divide by the pivot.  Then each thread takes trailing columns
// q in [j0+jb, j0+jb+KV): the panel's interchanges in order (rows c and
// ipiv[c], where column q still lies in row c's band), then the unit-lower
// triangle on the panel rows.
template <int ROWS>
__global__ __launch_bounds__(BPT) void k_band_panel(double *ab, int64_t LD, int KV, int kl, int64_t n, int64_t j0,
                                                    int *ipiv, int *info) {
  __shared__ double rowP[BNB], rowK[BNB];
  __shared__ double wv[BPT / 64];
  __shared__ int wi[BPT / 64], sp;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int jb = (int)(n - j0 < BNB ? n - j0 : BNB);
  const int PR = (int)((n < j0 + jb + kl ? n : j0 + jb + kl) - j0);
  double a[ROWS][BNB];
#pragma unroll
  for (int s = 0; s < ROWS; s++) {
    const int ii = t + s * BPT;
#pragma unroll
    for (int k = 0; k < BNB; k++) a[s][k] = (ii < PR && k < jb) ? ab[(j0 + k) * LD + KV + ii - k] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < BNB; k++) {
    if (k < jb) {
      const int hi = PR < k + kl + 1 ? PR : k + kl + 1;
      double bv = -1.0;
      int bi = INT_MAX;
#pragma unroll
      for (int s = 0; s < ROWS; s++) {
        const int ii = t + s * BPT;
        if (ii >= k && ii < hi) {
          const double v = fabs(a[s][k]);
          if (v > bv) {
            bv = v;
            bi = ii;
          }
        }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(bv, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > bv || (ov == bv && oi < bi)) {
          bv = ov;
          bi = oi;
        }
      }
      if (lane == 0) {
        wv[wave] = bv;
        wi[wave] = bi;
      }
      __syncthreads();
      if (t == 0) {
        double v = wv[0];
        int p = wi[0];
        for (int q = 1; q < BPT / 64; q++)
          if (wv[q] > v || (wv[q] == v && wi[q] < p)) {
            v = wv[q];
            p = wi[q];
          }
        sp = p;
        ipiv[j0 + k] = (int)(j0 + p);
      }
      __syncthreads();
      const int p = sp;
#pragma unroll
      for (int s = 0; s < ROWS; s++) {
        const int ii = t + s * BPT;
        if (ii == p)
#pragma unroll
          for (int q = 0; q < BNB; q++) rowP[q] = a[s][q];
        if (ii == k)
#pragma unroll
          for (int q = 0; q < BNB; q++) rowK[q] = a[s][q];
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < ROWS; s++) {
        const int ii = t + s * BPT;
        if (ii == k) {
#pragma unroll
          for (int q = 0; q < BNB; q++) a[s][q] = rowP[q];
        } else if (ii == p) {
#pragma unroll
          for (int q = 0; q < BNB; q++) a[s][q] = rowK[q];
        }
      }
      const double piv = rowP[k];
      if (piv == 0.0) {
        if (t == 0) atomicMin(info, (int)(j0 + k));
      } else {
#pragma unroll
        for (int s = 0; s < ROWS; s++) {
          const int ii = t + s * BPT;
          if (ii > k && ii < hi) {
            const double l = a[s][k] / piv;
            a[s][k] = l;
#pragma unroll
            for (int q = k + 1; q < BNB; q++) a[s][q] = a[s][q] - l * rowP[q];
          }
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int s = 0; s < ROWS; s++) {
    const int ii = t + s * BPT;
    if (ii < PR)
#pragma unroll
      for (int k = 0; k < BNB; k++)
        if (k < jb) ab[(j0 + k) * LD + KV + ii - k] = a[s][k];
  }
}

// The panel's interchanges and unit-lower triangle on the trailing columns
// q in [j0+jb, j0+jb+KV), one thread per column: rows c and ipiv[c] in
// order where column q still lies in row c's band, then the triangle on the
// panel rows.
__global__ __launch_bounds__(256) void k_band_trsm(double *ab, int64_t LD, int KV, int64_t n, int64_t j0,
                                                   const int *ipiv) {
  __shared__ double L11[BNB][BNB + 1];
  __shared__ int pvl[BNB];
  const int t = threadIdx.x;
  const int jb = (int)(n - j0 < BNB ? n - j0 : BNB);
  for (int e = t; e < BNB * BNB; e += 256) {
    const int r = e % BNB, k = e / BNB;
    L11[r][k] = (r > k && r < jb) ? ab[(j0 + k) * LD + KV + r - k] : 0.0;
  }
  if (t < BNB) pvl[t] = t < jb ? (int)(ipiv[j0 + t] - j0) : t;
  __syncthreads();
  const int64_t q0 = j0 + jb, q1 = n < j0 + jb + KV ? n : j0 + jb + KV;
  const int64_t q = q0 + (int64_t)blockIdx.x * 256 + t;
  if (q < q1) {
    double *col = ab + q * LD + KV;  // col[i - q] = A(i, q)
    for (int k = 0; k < jb; k++) {
      const int64_t c = j0 + k, p = j0 + pvl[k];
      if (p != c && q - c <= KV) {
        const double x = col[c - q];
        col[c - q] = col[p - q];
        col[p - q] = x;
      }
    }
    double uu[BNB];
#pragma unroll
    for (int k = 0; k < BNB; k++) uu[k] = (k < jb && q - (j0 + k) <= KV) ? col[j0 + k - q] : 0.0;
#pragma unroll
    for (int k = 0; k < BNB; k++)
#pragma unroll
      for (int k2 = k + 1; k2 < BNB; k2++) uu[k2] = uu[k2] - L11[k2][k] * uu[k];
#pragma unroll
    for (int k = 0; k < BNB; k++)
      if (k < jb && q - (j0 + k) <= KV) col[j0 + k - q] = uu[k];
  }
}

// A(i, q) -= sum_k L(i, j0+k) U(j0+k, q) for i in [j0+jb, j0+jb+kl),
// q in [j0+jb, j0+jb+KV): the window lies inside the band; U(j0+k, q) is
// zero where q lies beyond row j0+k's band.  64 x 64 tiles, each thread one
// row and 16 columns (coalesced along the rows of a column).
__global__ __launch_bounds__(256) void k_band_gemm(double *ab, int64_t LD, int KV, int kl, int64_t n, int64_t j0,
                                                   int jb) {
  __shared__ double U[BNB][65];
  const int64_t r0 = j0 + jb, r1 = n < j0 + jb + kl ? n : j0 + jb + kl;
  const int64_t q1 = n < j0 + jb + KV ? n : j0 + jb + KV;
  const int64_t tq = j0 + jb + (int64_t)blockIdx.y * 64, ti = r0 + (int64_t)blockIdx.x * 64;
  const int t = threadIdx.x;
  for (int e = t; e < BNB * 64; e += 256) {
    const int k = e % BNB, cq = e / BNB;
    const int64_t q = tq + cq;
    double v = 0;
    if (k < jb && q < q1 && q - (j0 + k) <= KV) v = ab[q * LD + KV + (j0 + k) - q];
    U[k][cq] = v;
  }
  __syncthreads();
  const int lane = t & 63, grp = t >> 6;
  const int64_t i = ti + lane;
  if (i >= r1) return;
  double l[BNB];
#pragma unroll
  for (int k = 0; k < BNB; k++) l[k] = k < jb ? ab[(j0 + k) * LD + KV + i - (j0 + k)] : 0.0;
  for (int cq = grp * 16; cq < grp * 16 + 16; cq++) {
    const int64_t q = tq + cq;
    if (q >= q1) break;
    double *cp = ab + q * LD + KV + i - q;
    double acc = *cp;
#pragma unroll
    for (int k = 0; k < BNB; k++) acc = fma(-l[k], U[k][cq], acc);
    *cp = acc;
  }
}

// ---------------------------------------------------------------- solves
// L y = P b, panel by panel: the panel's interchanges, its unit-lower
// triangle (one wavefront), then the rows below it; b[i] lives at
// win[i % BW] while it is active.
__global__ __launch_bounds__(BPT) void k_band_fwd(const double *ab, int64_t LD, int KV, int kl, int64_t n,
                                                  const int *ipiv, double *b) {
  __shared__ double win[BW];
  const int t = threadIdx.x, lane = t & 63;
  int64_t loaded = 0;
  for (int64_t j0 = 0; j0 < n; j0 += BNB) {
    const int jb = (int)(n - j0 < BNB ? n - j0 : BNB);
    const int64_t rmax = n < j0 + jb + kl ? n : j0 + jb + kl;
    for (int64_t i = loaded + t; i < rmax; i += BPT) win[i % BW] = b[i];
    loaded = rmax > loaded ? rmax : loaded;
    __syncthreads();
    if (t == 0)
      for (int k = 0; k < jb; k++) {
        const int64_t c = j0 + k, p = ipiv[c];
        if (p != c) {
          const double x = win[c % BW];
          win[c % BW] = win[p % BW];
          win[p % BW] = x;
        }
      }
    __syncthreads();
    if (t < 64) {
      double yr = lane < jb ? win[(j0 + lane) % BW] : 0.0;
      for (int k = 0; k < jb; k++) {
        const double yk = __shfl(yr, k);
        if (lane > k && lane < jb) yr -= ab[(j0 + k) * LD + KV + lane - k] * yk;
      }
      if (lane < jb) win[(j0 + lane) % BW] = yr;
    }
    __syncthreads();
    for (int64_t i = j0 + jb + t; i < rmax; i += BPT) {
      double acc = win[i % BW];
      for (int k = 0; k < jb; k++) acc -= ab[(j0 + k) * LD + KV + i - (j0 + k)] * win[(j0 + k) % BW];
      win[i % BW] = acc;
    }
    if (t < jb) b[j0 + t] = win[(j0 + t) % BW];
    __syncthreads();
  }
}

// U x = y, panels from the last: the panel's upper triangle (one
// wavefront), then the rows above it within KV.
__global__ __launch_bounds__(BPT) void k_band_bwd(const double *ab, int64_t LD, int KV, int64_t n, double *b) {
  __shared__ double win[BW];
  const int t = threadIdx.x, lane = t & 63;
  int64_t lo = n;  // b[lo .. n) are in the window
  for (int64_t j0 = ((n - 1) / BNB) * BNB; j0 >= 0; j0 -= BNB) {
    const int jb = (int)(n - j0 < BNB ? n - j0 : BNB);
    const int64_t rlo = j0 - KV > 0 ? j0 - KV : 0;
    for (int64_t i = rlo + t; i < lo; i += BPT) win[i % BW] = b[i];
    lo = rlo < lo ? rlo : lo;
    __syncthreads();
    if (t < 64) {
      double yr = lane < jb ? win[(j0 + lane) % BW] : 0.0;
      for (int k = jb - 1; k >= 0; k--) {
        if (lane == k) yr = yr / ab[(j0 + k) * LD + KV];
        const double xk = __shfl(yr, k);
        if (lane < k) yr -= ab[(j0 + k) * LD + KV + lane - k] * xk;
      }
      if (lane < jb) win[(j0 + lane) % BW] = yr;
    }
    __syncthreads();
    for (int64_t r = rlo + t; r < j0; r += BPT) {
      double acc = win[r % BW];
      for (int k = 0; k < jb; k++) {
        const int64_t q = j0 + k;
        if (q - r <= KV) acc -= ab[q * LD + KV + r - q] * win[q % BW];
      }
      win[r % BW] = acc;
    }
    if (t < jb) b[j0 + t] = win[(j0 + t) % BW];
    __syncthreads();
  }
}

__global__ void k_band_mul(int64_t n, const double *s, const double *x, double *y) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) y[k] = s[k] * x[k];
}

}  // namespace i2e

// ---------------------------------------------------------------- host
// band memory at most this (the 501^2 rows take ~27 GB)
constexpr double BAND_MAX_BYTES = 160e9;

int band_factor(i2e_ctx *c) {
  if (!c->band) c->band = new BandLU();
  BandLU &B = *c->band;
  const int X = c->X, Y = c->Y;
  const int64_t n = 3 * (int64_t)c->N;
  B.n = n;
  ST_OK(B.ecol.alloc((size_t)n * BNE));
  ST_OK(B.eval.alloc((size_t)n * BNE));
  ST_OK(B.ecnt.alloc(n));
  ST_OK(B.dims.alloc(4));
  ST_OK(B.rs.alloc(n));
  ST_OK(B.cs.alloc(n));
  ST_OK(B.rmx.alloc(n));
  ST_OK(B.cmx.alloc(n));
  ST_OK(B.w.alloc(n));
  ST_OK(B.ipiv.alloc(n));
  HIP_OK(hipMemsetAsync(B.dims.p, 0, 4 * sizeof(int), c->s));
  hipLaunchKernelGGL(k_band_rows, dim3((Y + 255) / 256, X), dim3(256), 0, c->s, c->sop(), c->bc, B.ecol.p, B.eval.p, B.ecnt.p,
                     B.dims.p);
  HIP_OK(hipGetLastError());
  int dims[4];
  HIP_OK(hipMemcpyAsync(dims, B.dims.p, sizeof(dims), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  if (dims[2]) {
    set_error("band LU: a Stokes row has more than %d coefficients", BNE);
    return I2E_ERR_UNSUPPORTED;
  }
  B.kl = dims[0];
  B.ku = dims[1];
  B.KV = B.kl + B.ku;
  B.LD = (int64_t)B.KV + B.kl + BNB;
  const double bytes = (double)n * (double)B.LD * sizeof(double);
  if (B.kl + BNB > BROWS * BPT || B.KV + BNB > BW || B.KV < BNB || bytes > BAND_MAX_BYTES) {
    set_error("band LU: bandwidths kl %d ku %d (%.1f GB) beyond what the band kernels take (kl + ku <= %d)", B.kl,
              B.ku, bytes / 1e9, BW - BNB);
    return I2E_ERR_UNSUPPORTED;
  }
  // Ruiz equilibration of the assembled rows
  const unsigned g = nblk(n, 256);
  hipLaunchKernelGGL(k_band_fill, dim3(1024), dim3(256), 0, c->s, n, 1.0, B.rs.p);
  hipLaunchKernelGGL(k_band_fill, dim3(1024), dim3(256), 0, c->s, n, 1.0, B.cs.p);
  for (int sw = 0; sw < BSWEEPS; sw++) {
    hipLaunchKernelGGL(k_band_zero_u64, dim3(1024), dim3(256), 0, c->s, n, B.cmx.p);
    hipLaunchKernelGGL(k_band_rowmax, dim3(g), dim3(256), 0, c->s, n, B.ecnt.p, B.eval.p, B.rmx.p);
    hipLaunchKernelGGL(k_band_colmax, dim3(g), dim3(256), 0, c->s, n, B.ecnt.p, B.ecol.p, B.eval.p, B.cmx.p);
    hipLaunchKernelGGL(k_band_scale, dim3(g), dim3(256), 0, c->s, n, B.ecnt.p, B.ecol.p, B.eval.p, B.rmx.p, B.cmx.p,
                       B.rs.p, B.cs.p);
  }
  ST_OK(B.ab.alloc((size_t)n * B.LD));
  hipLaunchKernelGGL(k_band_fill, dim3(4096), dim3(256), 0, c->s, n * B.LD, 0.0, B.ab.p);
  hipLaunchKernelGGL(k_band_scatter, dim3(g), dim3(256), 0, c->s, n, B.ecnt.p, B.ecol.p, B.eval.p, B.ab.p, B.LD,
                     B.KV);
  const int big = INT_MAX;
  HIP_OK(hipMemcpyAsync(B.dims.p + 3, &big, sizeof(int), hipMemcpyHostToDevice, c->s));
  for (int64_t j0 = 0; j0 < n; j0 += BNB) {
    const int jb = (int)(n - j0 < BNB ? n - j0 : BNB);
    if (B.kl + BNB <= BPT)
      hipLaunchKernelGGL(k_band_panel<1>, dim3(1), dim3(BPT), 0, c->s, B.ab.p, B.LD, B.KV, B.kl, n, j0, B.ipiv.p,
                         B.dims.p + 3);
    else
      hipLaunchKernelGGL(k_band_panel<2>, dim3(1), dim3(BPT), 0, c->s, B.ab.p, B.LD, B.KV, B.kl, n, j0, B.ipiv.p,
                         B.dims.p + 3);
    if (j0 + jb < n) {
      const int64_t nq0 = (n < j0 + jb + B.KV ? n : j0 + jb + B.KV) - (j0 + jb);
      hipLaunchKernelGGL(k_band_trsm, dim3(nblk(nq0, 256)), dim3(256), 0, c->s, B.ab.p, B.LD, B.KV, n, j0,
                         (const int *)B.ipiv.p);
      const int64_t nr = (n < j0 + jb + B.kl ? n : j0 + jb + B.kl) - (j0 + jb);
      const int64_t nq = (n < j0 + jb + B.KV ? n : j0 + jb + B.KV) - (j0 + jb);
      if (nr > 0 && nq > 0)
        hipLaunchKernelGGL(k_band_gemm, dim3(nblk(nr, 64), nblk(nq, 64)), dim3(256), 0, c->s, B.ab.p, B.LD, B.KV,
                           B.kl, n, j0, jb);
    }
  }
  HIP_OK(hipGetLastError());
  int info = 0;
  HIP_OK(hipMemcpyAsync(&info, B.dims.p + 3, sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  if (info != INT_MAX) {
    set_error("band LU: zero pivot in column %d of the equilibrated Stokes rows", info);
    return I2E_ERR_NOCONV;
  }
  return I2E_OK;
}

// z = C (LU)^-1 R r: the equilibrated factors applied to the unscaled rows
int band_apply(i2e_ctx *c, const double *r, double *z) {
  BandLU &B = *c->band;
  const int64_t n = B.n;
  const unsigned g = nblk(n, 256);
  hipLaunchKernelGGL(k_band_mul, dim3(g), dim3(256), 0, c->s, n, B.rs.p, r, B.w.p);
  hipLaunchKernelGGL(k_band_fwd, dim3(1), dim3(BPT), 0, c->s, B.ab.p, B.LD, B.KV, B.kl, n, B.ipiv.p, B.w.p);
  hipLaunchKernelGGL(k_band_bwd, dim3(1), dim3(BPT), 0, c->s, B.ab.p, B.LD, B.KV, n, B.w.p);
  hipLaunchKernelGGL(k_band_mul, dim3(g), dim3(256), 0, c->s, n, B.cs.p, B.w.p, z);
  HIP_OK(hipGetLastError());
  return I2E_OK;
}

int band_dims(const i2e_ctx *c, int *kl, int *ku, double *bytes) {
  if (!c->band) return I2E_ERR_ARG;
  *kl = c->band->kl;
  *ku = c->band->ku;
  *bytes = (double)c->band->n * (double)c->band->LD * sizeof(double);
  return I2E_OK;
}

extern "C" int i2e_band_info(i2e_ctx *c, int32_t *kl, int32_t *ku, double *bytes) {
  if (!c || !kl || !ku || !bytes) return I2E_ERR_ARG;
  if (!c->band) {
    set_error("band_info: no banded factorisation yet (stokes_direct = 1 and a Stokes solve)");
    return I2E_ERR_ARG;
  }
  return band_dims(c, kl, ku, bytes);
}
