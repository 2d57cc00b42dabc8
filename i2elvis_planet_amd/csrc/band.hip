// band.hip -- the Stokes rows assembled into band storage and factorised by
// LU with partial pivoting inside the band: the FGMRES preconditioner of
// stokes_direct = 1 and 2.  The reference solves every Stokes system directly
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
//      partial pivoting (the panel's multipliers end in the permuted order
//      of its interchanges, as in LAPACK's dgetrf; band storage keeps them
//      in BNB extra rows), a second kernel applies the interchanges and the
//      unit-lower triangle to the panel rows of the trailing columns, and
//      the trailing band window (kl rows x KV columns) takes the rank-BNB
//      update, the next panel's columns first when the window is wide;
//   5. the solves: one workgroup each, the active part of the vector in an
//      LDS window, panel by panel, with the row moves and pivot reciprocals
//      the factorisation recorded.
#include <climits>

#include "ctx.hpp"
#include "ops.hpp"
#include "rows.hpp"

namespace i2e {

constexpr int BNE = 40;    // assembled entries per row at most
constexpr int BNB = 16;    // panel width
constexpr int BPT = 1024;  // panel and solve workgroup
constexpr int BPMAX = 2 * 768;  // panel rows at most: kl + BNB <= BPMAX
constexpr int BW = 4096;   // LDS window of the solves (doubles): kl + BNB, KV + BNB <= BW
constexpr int BSWEEPS = 10;

struct BandLU {
  int64_t n = 0;  // unknowns (3 N)
  int kl = 0, ku = 0, KV = 0;
  int64_t LD = 0;
  DBuf<double> ab, rs, cs, rmx, w, w2, eval, dinv;
  DBuf<unsigned long long> cmx;
  DBuf<int> ipiv, ecol, ecnt, dims, mv;
  hipEvent_t ev_t = nullptr, ev_b = nullptr;  // trailing interchanges done; the rest of an update done
};

void band_release(BandLU *b) {
  if (!b) return;
  b->ab.release();
  b->rs.release();
  b->cs.release();
  b->rmx.release();
  b->w.release();
  b->w2.release();
  b->eval.release();
  b->cmx.release();
  b->ipiv.release();
  b->dinv.release();
  b->mv.release();
  if (b->ev_t) (void)hipEventDestroy(b->ev_t);
  if (b->ev_b) (void)hipEventDestroy(b->ev_b);
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
// Pivot search key of a candidate row: |v|'s bits with the low 12 mantissa
// bits replaced by 4095 - row, so that the largest key is the largest
// magnitude (to 2^-40 relative) and, among those, the first row; one
// 64-bit max then finds both.
__device__ __forceinline__ unsigned long long piv_key(double v, int row) {
  return ((unsigned long long)__double_as_longlong(fabs(v)) & ~0xFFFull) | (unsigned long long)(0xFFF - row);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long dpp_max_u64(unsigned long long k) {
  const unsigned long long o = (unsigned long long)__double_as_longlong(dpp_get<CTRL, ROW_MASK>(__longlong_as_double((long long)k)));
  return o > k ? o : k;
}
// maximum over the wavefront's 64 lanes (all active), to every lane
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long k) {
  k = dpp_max_u64<0xB1, 0xF>(k);
  k = dpp_max_u64<0x4E, 0xF>(k);
  k = dpp_max_u64<0x141, 0xF>(k);
  k = dpp_max_u64<0x140, 0xF>(k);
  k = dpp_max_u64<0x142, 0xA>(k);
  k = dpp_max_u64<0x143, 0xC>(k);
  const int lo = __builtin_amdgcn_readlane((int)k, 63), hi = __builtin_amdgcn_readlane((int)(k >> 32), 63);
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// Panel columns j0 .. j0+jb-1, rows j0 .. j0+jb+kl-1 (below that the panel
// columns are zero).  Thread t holds panel rows t + s NT (s < ROWS) in
// registers and keeps them: a pivot choice only renames rows (each row
// keeps its position in a register; the chosen row and the row at the
// pivot's position exchange positions, as the interchanges of partial
// pivoting would move them).  Each column is one
// pass of a rolled loop, so the code stays in the instruction cache: the
// column in hand is a[0] of every row, the rows shift left after each
// column, and the value a row leaves for the column (its multiplier, or its
// U entry once it has been chosen as pivot) shifts into out[BNB-1].  The
// pivot is the largest magnitude among rows not yet chosen (piv_key; per
// wavefront by DPP, then across wavefronts through LDS); the chosen row's
// values pass through LDS to the rank-1 update of the others.  At the end
// every row is stored at its final position.
// Besides the factors it leaves, per panel, the pivots' reciprocals
// (dinv, for the backward solve) and the row moves its interchanges make,
// for the forward solve: mv[64 panel + 2 s] = position, mv[.. + 1] = the
// position whose value lands there (relative to j0; -1: unused), over the
// panel rows and the rows below named by an interchange.
template <int ROWS, int NT>
__global__ __launch_bounds__(NT) void k_band_panel(double *ab, int64_t LD, int KV, int kl, int64_t n, int64_t j0,
                                                   int *ipiv, int *info, double *dinv, int *mv) {
  static_assert(BPMAX < 4095, "pivot keys hold 12 bits of row");
  static_assert(ROWS <= 2, "two rows per thread at most");
  __shared__ double rowP[BNB + 1];  // the pivot row and, last, 1 / pivot
  __shared__ unsigned long long wk[NT / 64];
  __shared__ int ppos, nmv, mpos[2 * BNB], msrc[2 * BNB];
  // the launch sizes the block to the panel's rows (whole wavefronts)
  const int nt = blockDim.x, NW = nt / 64;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int jb = (int)(n - j0 < BNB ? n - j0 : BNB);
  const int PR = (int)((n < j0 + jb + kl ? n : j0 + jb + kl) - j0);
  // one array per held row and value (a 2-D array is not kept in registers)
  double a0[BNB], a1[BNB], o0[BNB], o1[BNB];
  int d0 = 0, d1 = 0;             // chosen as pivot already
  int p0 = t, p1 = t + nt;        // current positions
  auto rows = [&](auto f) {
    f(a0, o0, d0, t);
    if constexpr (ROWS > 1) f(a1, o1, d1, t + nt);
  };
  auto posof = [&](int ii) -> int & { return ii == t ? p0 : p1; };
  if (t == 0) nmv = 0;
  rows([&](double(&ar)[BNB], double(&orow)[BNB], int &, int ii) {
#pragma unroll
    for (int c = 0; c < BNB; c++) {
      ar[c] = (ii < PR && c < jb) ? ab[(j0 + c) * LD + KV + ii - c] : 0.0;
      orow[c] = 0.0;
    }
  });
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < BNB; k++) {
    if (k < jb) {
      unsigned long long key = 0;
      rows([&](double(&ar)[BNB], double(&)[BNB], int &done, int ii) {
        if (!done && ii < PR) {
          const unsigned long long kk = piv_key(ar[0], ii);
          key = kk > key ? kk : key;
        }
      });
      key = wave_max_u64(key);
      if (lane == 0) wk[wave] = key;
      __syncthreads();
      for (int w = 0; w < NW; w++) key = wk[w] > key ? wk[w] : key;
      const int r = 0xFFF - (int)(key & 0xFFF);  // the chosen row
      rows([&](double(&ar)[BNB], double(&)[BNB], int &, int ii) {
        if (ii == r) {
#pragma unroll
          for (int c = 0; c < BNB; c++) rowP[c] = ar[c];
          rowP[BNB] = ar[0] != 0.0 ? 1.0 / ar[0] : 0.0;
          ppos = posof(ii);
          ipiv[j0 + k] = (int)(j0 + posof(ii));
        }
      });
      __syncthreads();
      {
        // the chosen row moves to position k, the row there to its place
        const int pk = ppos;
        rows([&](double(&)[BNB], double(&)[BNB], int &, int ii) {
          int &ps = posof(ii);
          if (ii == r)
            ps = k;
          else if (ps == k)
            ps = pk;
        });
      }
      // multipliers scaled by the pivot's reciprocal, as LAPACK's dgetf2 does
      const double rpiv = rowP[BNB];
      if (t == 0) {
        if (rpiv == 0.0) atomicMin(info, (int)(j0 + k));
        dinv[j0 + k] = rpiv;
      }
      rows([&](double(&ar)[BNB], double(&orow)[BNB], int &done, int ii) {
        double enter = ar[0];
        if (ii == r) {
          done = 1;
        } else if (!done) {
          const double l = ar[0] * rpiv;
          enter = l;
#pragma unroll
          for (int c = 1; c < BNB; c++) ar[c] = ar[c] - l * rowP[c];
        }
#pragma unroll
        for (int c = 0; c + 1 < BNB; c++) {
          ar[c] = ar[c + 1];
          orow[c] = orow[c + 1];
        }
        ar[BNB - 1] = 0.0;
        orow[BNB - 1] = enter;
      });
    } else {
      // a last panel narrower than BNB: shift on so that out[c] is column c
      rows([&](double(&)[BNB], double(&orow)[BNB], int &, int) {
#pragma unroll
        for (int c = 0; c + 1 < BNB; c++) orow[c] = orow[c + 1];
        orow[BNB - 1] = 0.0;
      });
    }
  }
  // row moves (at most two per interchange): position, and the position
  // the row there started from
  rows([&](double(&)[BNB], double(&)[BNB], int &, int ii) {
    const int pos = posof(ii);
    if (ii < PR && pos != ii) {
      const int e = atomicAdd(&nmv, 1);
      mpos[e] = pos;
      msrc[e] = ii;
    }
  });
  __syncthreads();
  int *m = mv + (j0 / BNB) * 4 * BNB;
  if (t < 2 * BNB) {
    m[2 * t] = t < nmv ? mpos[t] : -1;
    m[2 * t + 1] = t < nmv ? msrc[t] : -1;
  }
  rows([&](double(&)[BNB], double(&orow)[BNB], int &, int ii) {
    if (ii < PR) {
      const int pos = posof(ii);
#pragma unroll
      for (int c = 0; c < BNB; c++)
        if (c < jb) ab[(j0 + c) * LD + KV + pos - c] = orow[c];
    }
  });
}

// Where the panel's interchanges put rows: slot k < BNB is panel row j0+k,
// slot BNB + e the row below the panel that interchange e names first
// (xr[e]; -1 where interchange e names a panel row or a row named before);
// the row of interchange k sits in slot sl[k] (>= k).  Thread k < jb
// fills entry k from the panel's pivots pv[] (in LDS).
__device__ __forceinline__ void panel_slot(const int *pv, int64_t j0, int jb, int k, int *sl, int64_t *xr) {
  const int64_t p = pv[k];
  if (p < j0 + jb) {
    sl[k] = (int)(p - j0);
    xr[k] = -1;
    return;
  }
  int e = 0;
  while (pv[e] != p) e++;  // the first interchange naming p (e <= k)
  sl[k] = BNB + e;
  xr[k] = e == k ? p : -1;
}

// The panel's interchanges and unit-lower triangle on the trailing columns
// q in [j0+jb, j0+jb+KV): 16 lanes per column, lane k loading panel row
// j0+k (16 consecutive doubles per column) and the row below the panel that
// interchange k names first; the interchanges, composed once per block
// into one permutation of the 32 slots, move the values through LDS, and
// the triangle runs across the 16 lanes by shuffles.  (Where column q lies
// beyond a panel row's band, that row and the row it exchanges with both
// hold zero there.)
__global__ __launch_bounds__(256) void k_band_trsm(double *ab, int64_t LD, int KV, int64_t n, int64_t j0,
                                                   const int *ipiv) {
  __shared__ double L11[BNB][BNB + 1];
  __shared__ double vs[256 / BNB][2 * BNB + 1];
  __shared__ int sl[BNB], pv[BNB], perm[2 * BNB];
  __shared__ int64_t xr[BNB];
  const int t = threadIdx.x, cq = t / BNB, k = t % BNB;
  const int jb = (int)(n - j0 < BNB ? n - j0 : BNB);
  {
    const int r = t % BNB, c = t / BNB;
    L11[r][c] = (r > c && r < jb) ? ab[(j0 + c) * LD + KV + r - c] : 0.0;
  }
  if (t < jb) pv[t] = ipiv[j0 + t];
  __syncthreads();
  if (t < jb) panel_slot(pv, j0, jb, t, sl, xr);
  if (t >= jb && t < BNB) xr[t] = -1;
  __syncthreads();
  if (t == 0) {
    // slot s ends up holding the value first in slot perm[s]
    for (int e = 0; e < 2 * BNB; e++) perm[e] = e;
    for (int c = 0; c < jb; c++) {
      const int s2 = sl[c];
      const int x = perm[c];
      perm[c] = perm[s2];
      perm[s2] = x;
    }
  }
  const int64_t q1 = n < j0 + jb + KV ? n : j0 + jb + KV;
  const int64_t q = j0 + jb + (int64_t)blockIdx.x * (256 / BNB) + cq;
  const bool act = q < q1;
  double *col = ab + (act ? q : 0) * LD + KV;  // col[i - q] = A(i, q)
  const bool inb = act && k < jb && q - (j0 + k) <= KV;
  const int64_t xk = xr[k];
  vs[cq][k] = inb ? col[j0 + k - q] : 0.0;
  vs[cq][BNB + k] = (act && xk >= 0) ? col[xk - q] : 0.0;
  __syncthreads();
  double u = vs[cq][perm[k]];
  const double ue = vs[cq][perm[BNB + k]];
#pragma unroll
  for (int c = 0; c < BNB; c++) {
    const double uc = __shfl(u, c, BNB);
    if (k > c) u = u - L11[k][c] * uc;
  }
  if (inb) col[j0 + k - q] = u;
  if (act && xk >= 0) col[xk - q] = ue;
}

// A(i, q) -= sum_k L(i, j0+k) U(j0+k, q) for i in [j0+jb, j0+jb+kl),
// q in [qa, qb) within [j0+jb, j0+jb+KV) (the factorisation runs the next
// panel's columns first, the rest beside the next panel on a second
// stream): the window lies inside the band; U(j0+k, q) is
// zero where q lies beyond row j0+k's band.  64 x 64 tiles, each thread one
// row and 16 columns (coalesced along the rows of a column), the 16 loads
// issued together.
__global__ __launch_bounds__(256) void k_band_gemm(double *ab, int64_t LD, int KV, int kl, int64_t n, int64_t j0,
                                                   int jb, int64_t qa, int64_t qb) {
  __shared__ double U[64][BNB + 1];  // U[cq][k] = U(j0+k, tq+cq)
  const int64_t r0 = j0 + jb, r1 = n < j0 + jb + kl ? n : j0 + jb + kl;
  const int64_t q1 = qb;
  const int64_t tq = qa + (int64_t)blockIdx.y * 64, ti = r0 + (int64_t)blockIdx.x * 64;
  const int t = threadIdx.x;
  for (int e = t; e < BNB * 64; e += 256) {
    const int k = e % BNB, cq = e / BNB;
    const int64_t q = tq + cq;
    double v = 0;
    if (k < jb && q < q1 && q - (j0 + k) <= KV) v = ab[q * LD + KV + (j0 + k) - q];
    U[cq][k] = v;
  }
  __syncthreads();
  const int lane = t & 63, grp = t >> 6;
  const int64_t i = ti + lane;
  if (i >= r1) return;
  // A(i, q) for consecutive q lies LD - 1 apart
  const int64_t st = LD - 1, qc = tq + grp * 16;
  const int nc = (int)(q1 - qc < 16 ? q1 - qc : 16);
  double *base = ab + qc * st + KV + i;
  const double *lb = ab + j0 * st + KV + i;
  double l[BNB];
#pragma unroll
  for (int k = 0; k < BNB; k++) l[k] = k < jb ? lb[k * st] : 0.0;
  double cv[16];
#pragma unroll
  for (int c = 0; c < 16; c++) cv[c] = c < nc ? base[c * st] : 0.0;
#pragma unroll
  for (int c = 0; c < 16; c++) {
    double acc = cv[c];
#pragma unroll
    for (int k = 0; k < BNB; k++) acc = fma(-l[k], U[grp * 16 + c][k], acc);
    if (c < nc) base[c * st] = acc;
  }
}

// ---------------------------------------------------------------- solves
// L y = P b, panel by panel: b[i] lives at win[i % BW] while it is active.
// Per panel: the new entries and the panel's unit-lower triangle come into
// LDS together; wavefront 0 makes the panel's row moves (all reads before
// all writes), then solves the triangle by lane shuffles; then every
// thread updates rows below the panel, its 16 loads issued together.
__global__ __launch_bounds__(BPT) void k_band_fwd(const double *ab, int64_t LD, int KV, int kl, int64_t n,
                                                  const int *mv, double *b) {
  __shared__ double win[BW], tri[BNB][BNB + 1];
  const int t = threadIdx.x, lane = t & 63;
  int64_t loaded = 0;
  for (int64_t j0 = 0; j0 < n; j0 += BNB) {
    const int jb = (int)(n - j0 < BNB ? n - j0 : BNB);
    const int64_t rmax = n < j0 + jb + kl ? n : j0 + jb + kl;
    const int *m = mv + (j0 / BNB) * 4 * BNB;
    const int pos = t < 2 * BNB ? m[2 * t] : -1, src = t < 2 * BNB ? m[2 * t + 1] : -1;
    // this thread's rows below the panel (kl <= 2 BPT), their multipliers
    // loaded now so that the loads overlap the triangle
    const int64_t i0 = j0 + jb + t, i1 = i0 + BPT;
    double l0[BNB], l1[BNB];
#pragma unroll
    for (int k = 0; k < BNB; k++) {
      l0[k] = (k < jb && i0 < rmax) ? ab[(j0 + k) * LD + KV + i0 - (j0 + k)] : 0.0;
      l1[k] = (k < jb && i1 < rmax) ? ab[(j0 + k) * LD + KV + i1 - (j0 + k)] : 0.0;
    }
    for (int64_t i = loaded + t; i < rmax; i += BPT) win[i % BW] = b[i];
    loaded = rmax > loaded ? rmax : loaded;
    if (t < BNB * BNB) {
      const int r = t % BNB, k = t / BNB;
      tri[r][k] = (r > k && r < jb) ? ab[(j0 + k) * LD + KV + r - k] : 0.0;
    }
    __syncthreads();
    if (t < 64) {
      const double v = pos >= 0 ? win[(j0 + src) % BW] : 0.0;
      if (pos >= 0) win[(j0 + pos) % BW] = v;
      double yr = lane < jb ? win[(j0 + lane) % BW] : 0.0;
      for (int k = 0; k < jb; k++) {
        const double yk = __shfl(yr, k);
        if (lane > k && lane < jb) yr -= tri[lane][k] * yk;
      }
      if (lane < jb) win[(j0 + lane) % BW] = yr;
    }
    __syncthreads();
    if (i0 < rmax) {
      double acc = win[i0 % BW];
#pragma unroll
      for (int k = 0; k < BNB; k++)
        if (k < jb) acc -= l0[k] * win[(j0 + k) % BW];
      win[i0 % BW] = acc;
    }
    if (i1 < rmax) {
      double acc = win[i1 % BW];
#pragma unroll
      for (int k = 0; k < BNB; k++)
        if (k < jb) acc -= l1[k] * win[(j0 + k) % BW];
      win[i1 % BW] = acc;
    }
    if (t < jb) b[j0 + t] = win[(j0 + t) % BW];
    __syncthreads();
  }
}

// U x = y, panels from the last: the panel's upper triangle comes into LDS
// with the new entries, wavefront 0 solves it (multiplying by the pivots'
// reciprocals), then the rows above within KV take the panel's columns,
// 16 loads a row issued together.
__global__ __launch_bounds__(BPT) void k_band_bwd(const double *ab, int64_t LD, int KV, int64_t n,
                                                  const double *dinv, double *b) {
  __shared__ double win[BW], tri[BNB][BNB + 1];
  const int t = threadIdx.x, lane = t & 63;
  int64_t lo = n;  // b[lo .. n) are in the window
  for (int64_t j0 = ((n - 1) / BNB) * BNB; j0 >= 0; j0 -= BNB) {
    const int jb = (int)(n - j0 < BNB ? n - j0 : BNB);
    const int64_t rlo = j0 - KV > 0 ? j0 - KV : 0;
    for (int64_t i = rlo + t; i < lo; i += BPT) win[i % BW] = b[i];
    lo = rlo < lo ? rlo : lo;
    if (t < BNB * BNB) {
      const int r = t % BNB, k = t / BNB;
      tri[r][k] = (r < k && k < jb) ? ab[(j0 + k) * LD + KV + r - k] : 0.0;
    }
    const double di = (t < 64 && lane < jb) ? dinv[j0 + lane] : 0.0;
    // this thread's first two rows above the panel, their U entries loaded
    // now so that the loads overlap the triangle
    const int64_t r0 = rlo + t, r1 = r0 + BPT;
    double u0[BNB], u1[BNB];
#pragma unroll
    for (int k = 0; k < BNB; k++) {
      const int64_t q = j0 + k;
      u0[k] = (k < jb && r0 < j0 && q - r0 <= KV) ? ab[q * LD + KV + r0 - q] : 0.0;
      u1[k] = (k < jb && r1 < j0 && q - r1 <= KV) ? ab[q * LD + KV + r1 - q] : 0.0;
    }
    __syncthreads();
    if (t < 64) {
      double yr = lane < jb ? win[(j0 + lane) % BW] : 0.0;
      for (int k = jb - 1; k >= 0; k--) {
        if (lane == k) yr = yr * di;
        const double xk = __shfl(yr, k);
        if (lane < k) yr -= tri[lane][k] * xk;
      }
      if (lane < jb) win[(j0 + lane) % BW] = yr;
    }
    __syncthreads();
    if (r0 < j0) {
      double acc = win[r0 % BW];
#pragma unroll
      for (int k = 0; k < BNB; k++)
        if (k < jb) acc -= u0[k] * win[(j0 + k) % BW];
      win[r0 % BW] = acc;
    }
    if (r1 < j0) {
      double acc = win[r1 % BW];
#pragma unroll
      for (int k = 0; k < BNB; k++)
        if (k < jb) acc -= u1[k] * win[(j0 + k) % BW];
      win[r1 % BW] = acc;
    }
    for (int64_t r = rlo + t + 2 * BPT; r < j0; r += BPT) {
      double uv[BNB];
#pragma unroll
      for (int k = 0; k < BNB; k++) {
        const int64_t q = j0 + k;
        uv[k] = (k < jb && q - r <= KV) ? ab[q * LD + KV + r - q] : 0.0;
      }
      double acc = win[r % BW];
#pragma unroll
      for (int k = 0; k < BNB; k++)
        if (k < jb) acc -= uv[k] * win[(j0 + k) % BW];
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
__global__ void k_band_div(int64_t n, const double *s, double *y) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) y[k] = y[k] / s[k];
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
  ST_OK(B.w2.alloc(n));
  ST_OK(B.ipiv.alloc(n));
  ST_OK(B.dinv.alloc(n));
  ST_OK(B.mv.alloc((size_t)((n + BNB - 1) / BNB) * 4 * BNB));
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
  if (B.kl + BNB > BPMAX || B.KV + BNB > BW || B.KV < BNB || bytes > BAND_MAX_BYTES) {
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
  // Look-ahead: each panel's update runs the next panel's 16 columns on the
  // main stream, which then factors the next panel, while the rest of the
  // window is updated on the side stream; the next panel's trailing
  // interchanges wait for that rest (they touch the same columns).  Only
  // for wide windows: at 501^2 (KV = 3010) it took the solve from 4.67 to
  // 4.37 s, at 201^2 (KV = 1210) the two events per panel cost more than
  // the overlap gave (0.476 -> 0.503 s).
  const bool ahead = B.KV > 2048 && c->s2;
  if (!B.ev_t) HIP_OK(hipEventCreateWithFlags(&B.ev_t, hipEventDisableTiming));
  if (!B.ev_b) HIP_OK(hipEventCreateWithFlags(&B.ev_b, hipEventDisableTiming));
  hipStream_t side = c->s2 ? c->s2 : c->s;
  bool pending = false;
  for (int64_t j0 = 0; j0 < n; j0 += BNB) {
    const int jb = (int)(n - j0 < BNB ? n - j0 : BNB);
    // panel rows per thread: 1 at 1024 threads, else 2 at 768 (the
    // registers 768 threads leave hold 64 doubles a thread without spills)
    // (the block: the panel's rows rounded up to whole wavefronts)
    if (B.kl + BNB <= 1024)
      hipLaunchKernelGGL((k_band_panel<1, 1024>), dim3(1), dim3(((B.kl + BNB + 63) / 64) * 64), 0, c->s, B.ab.p,
                         B.LD, B.KV, B.kl, n, j0, B.ipiv.p, B.dims.p + 3, B.dinv.p, B.mv.p);
    else
      hipLaunchKernelGGL((k_band_panel<2, 768>), dim3(1), dim3(((B.kl + BNB + 127) / 128) * 64), 0, c->s, B.ab.p,
                         B.LD, B.KV, B.kl, n, j0, B.ipiv.p, B.dims.p + 3, B.dinv.p, B.mv.p);
    if (j0 + jb < n) {
      if (pending) HIP_OK(hipStreamWaitEvent(c->s, B.ev_b, 0));
      pending = false;
      const int64_t q0 = j0 + jb, q1 = n < j0 + jb + B.KV ? n : j0 + jb + B.KV;
      hipLaunchKernelGGL(k_band_trsm, dim3(nblk(q1 - q0, 256 / BNB)), dim3(256), 0, c->s, B.ab.p, B.LD, B.KV, n,
                         j0, (const int *)B.ipiv.p);
      const int64_t nr = (n < j0 + jb + B.kl ? n : j0 + jb + B.kl) - q0;
      const int64_t qs = !ahead ? q1 : q1 < q0 + BNB ? q1 : q0 + BNB;
      if (nr > 0) {
        hipLaunchKernelGGL(k_band_gemm, dim3(nblk(nr, 64), nblk(qs - q0, 64)), dim3(256), 0, c->s, B.ab.p, B.LD,
                           B.KV, B.kl, n, j0, jb, q0, qs);
        if (qs < q1) {
          HIP_OK(hipEventRecord(B.ev_t, c->s));
          HIP_OK(hipStreamWaitEvent(side, B.ev_t, 0));
          hipLaunchKernelGGL(k_band_gemm, dim3(nblk(nr, 64), nblk(q1 - qs, 64)), dim3(256), 0, side, B.ab.p, B.LD,
                             B.KV, B.kl, n, j0, jb, qs, q1);
          HIP_OK(hipEventRecord(B.ev_b, side));
          pending = true;
        }
      }
    }
  }
  if (pending) HIP_OK(hipStreamWaitEvent(c->s, B.ev_b, 0));
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

// unscaled rows: z = C (LU)^-1 R r, the equilibrated factors applied to the
// assembled system; equilibrated rows (scaled): z = (LU)^-1 r
int band_apply(i2e_ctx *c, const double *r, double *z, bool scaled) {
  BandLU &B = *c->band;
  const int64_t n = B.n;
  const unsigned g = nblk(n, 256);
  if (scaled)
    HIP_OK(hipMemcpyAsync(B.w.p, r, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  else
    hipLaunchKernelGGL(k_band_mul, dim3(g), dim3(256), 0, c->s, n, B.rs.p, r, B.w.p);
  hipLaunchKernelGGL(k_band_fwd, dim3(1), dim3(BPT), 0, c->s, B.ab.p, B.LD, B.KV, B.kl, n, (const int *)B.mv.p,
                     B.w.p);
  hipLaunchKernelGGL(k_band_bwd, dim3(1), dim3(BPT), 0, c->s, B.ab.p, B.LD, B.KV, n, (const double *)B.dinv.p,
                     B.w.p);
  if (scaled)
    HIP_OK(hipMemcpyAsync(z, B.w.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  else
    hipLaunchKernelGGL(k_band_mul, dim3(g), dim3(256), 0, c->s, n, B.cs.p, B.w.p, z);
  HIP_OK(hipGetLastError());
  return I2E_OK;
}

// The equilibrated system R A C y = R b (x = C y): b <- R b, x <- x / C
// before the solve, x <- C y after it; y -> R A C y through the
// matrix-free rows
int band_scale_system(i2e_ctx *c, double *b, double *x) {
  BandLU &B = *c->band;
  const unsigned g = nblk(B.n, 256);
  hipLaunchKernelGGL(k_band_mul, dim3(g), dim3(256), 0, c->s, B.n, B.rs.p, b, b);
  hipLaunchKernelGGL(k_band_div, dim3(g), dim3(256), 0, c->s, B.n, B.cs.p, x);
  HIP_OK(hipGetLastError());
  return I2E_OK;
}
int band_unscale(i2e_ctx *c, double *x) {
  BandLU &B = *c->band;
  hipLaunchKernelGGL(k_band_mul, dim3(nblk(B.n, 256)), dim3(256), 0, c->s, B.n, B.cs.p, x, x);
  HIP_OK(hipGetLastError());
  return I2E_OK;
}
int band_op(i2e_ctx *c, const double *y, double *out) {
  BandLU &B = *c->band;
  const unsigned g = nblk(B.n, 256);
  hipLaunchKernelGGL(k_band_mul, dim3(g), dim3(256), 0, c->s, B.n, B.cs.p, y, B.w2.p);
  hipLaunchKernelGGL(k_stokes_apply, dim3((c->Y + 255) / 256, c->X), dim3(256), 0, c->s, c->sop(), c->bc,
                     (const double *)B.w2.p, out);
  hipLaunchKernelGGL(k_band_mul, dim3(g), dim3(256), 0, c->s, B.n, B.rs.p, out, out);
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
