// mg.hpp -- geometric multigrid preconditioners (velocity block of Stokes,
// scalar heat/gravity), templated on the arithmetic type.
//
// Coarse grids keep every second grid line of the finer grid plus the last
// one (so odd cell counts such as 125 work, at the price of one narrow last
// cell); transfers are bilinear interpolation and its normalised transpose.
//
// Velocity block: the finest level is matrix-free (the exact rows) with
// red-black Gauss-Seidel per component; every coarser level is the Galerkin
// product R A P stored as an explicit 5x5-per-component stencil, computed on
// the device each solve, and smoothed by per-colour Jacobi (read one buffer,
// write the other: deterministic although the stencil couples same-colour
// points).  Rediscretised coarse operators with averaged viscosities were
// tried first: near viscosity jumps they are far softer (or, with max
// averaging, stiffer) than R A P and the V-cycle diverges or stalls.  Each
// smoothing pass ends with an exact re-solve of the pointwise boundary rows.
// The coarsest grid is solved with an explicit dense inverse.
//
// Scalar operators (heat, gravity): rediscretised with restricted
// coefficients and red-black Gauss-Seidel.
#pragma once
#include <vector>

#include "ops.hpp"

namespace i2e {

// device buffer
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    if (count <= n && p) return I2E_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (!count) return I2E_OK;
    HIP_OK(hipMalloc((void **)&p, count * sizeof(T)));
    n = count;
    return I2E_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// 1D interpolation map fine -> coarse coordinate sets
// restriction of one coarse index gathers at most RW consecutive fine indices
constexpr int RW = 4;

struct Map1 {
  int nf = 0, nc = 0;
  DBuf<int> i0;      // [nf]
  DBuf<double> w;    // [nf] weight of i0+1
  DBuf<int> rb, re;  // [nc] fine index range touching coarse index
  DBuf<double> rw;   // [nc*RW] normalised restriction weights of fine indices rb.. (0 past re)
  std::vector<int> hrb, hre;  // host copies of rb, re
  int build(const std::vector<double> &cf, const std::vector<double> &cc);
  // widest fine index range that tiles of t consecutive coarse indices gather from
  int tile_span(int t) const;
};

struct MapView {
  const int *i0;
  const double *w;
  const int *rb, *re;
  const double *rw;
};
inline MapView view(const Map1 &m) { return MapView{m.i0.p, m.w.p, m.rb.p, m.re.p, m.rw.p}; }

// level geometry shared by all solvers on one grid hierarchy
struct GeoLevel {
  int X, Y;
  std::vector<double> hx, hy;   // host copies of the lines
  DBuf<double> gx, gy;          // device lines
  std::vector<int> fi, fj;      // fine line index of each coarse line (to the next finer level)
};

// ---------------------------------------------------------------- velocity MG
template <class T>
struct VLev {
  SOp<T> o;
  const int8_t *kx, *ky;  // 0 stencil row, 1 BC row, 2 ghost row
  const int *nbx, *nby;   // BC neighbour (same component), -1 none
  const T *kkx, *kky;     // BC coefficient
};

template <class T>
struct VelLevel {
  int X, Y, N;
  DBuf<T> nu, nd, ro, gp;
  DBuf<int8_t> kx, ky;
  DBuf<int> nbx, nby;
  DBuf<T> kkx, kky;
  DBuf<T> dx, dy;           // diagonals
  DBuf<T> ux, uy, fx, fy, rx, ry;
  DBuf<T> ux2, uy2;         // second buffers of the coarse-level smoother
  DBuf<T> S;                // explicit stencil: level 0 probed (3x3), coarse levels Galerkin (5x5)
  int W = 2;                // stencil half-width
  Map1 mxn, mxm, myn, mym;  // maps to the next coarser level: x nodes/mids, y nodes/mids
  std::vector<int8_t> hkx, hky;
  std::vector<int> hnbx, hnby;
  std::vector<T> hkkx, hkky;
};

// (width along j, height along i) of tile shape candidate k
template <class T>
void tile_shape(int k, int *sx, int *sy);

template <class T>
struct VelMG {
  std::vector<VelLevel<T>> L;
  const std::vector<GeoLevel> *geo = nullptr;
  int nc = 0;               // coarsest unknown count (2 N)
  DBuf<double> dense;       // n x 2n work
  DBuf<double> ainv;        // n x n, FP64 in every precision (the coarsest solve stays exact)
  DBuf<double> ainvT;       // its transpose (column c at c*n: one thread per row reads coalesced)
  std::vector<DBuf<int>> gfi, gfj;  // fine line index of each coarse line, per level
  DBuf<int> overflow;       // Galerkin entries that fell outside the 5x5 window (must stay 0)
  double mu = 0, dt = 0;
  int nsmooth = 2;
  double omega = 1.0;       // damping of the smoothers' colour steps (exact rows undamped)
  int nsmooth_c = 0;        // sweeps on the coarser levels (0: nsmooth)
  int stencil0 = 0;         // finest level smoothed with its probed 3x3 stencil (else matrix-free rows)
  const double *S0d = nullptr;  // that stencil in FP64 (from the FP64 hierarchy) for FP32 iterates
  int diff0 = 0;            // FP32 hierarchy: finest stencil stored in difference form (refresh_from)
  int coop = 0;             // FP32: the levels from coop_level() down in one cooperative launch
  KTimer *kt = nullptr;     // optional per-launch timing of the smoothers
  int setup(const std::vector<GeoLevel> &g, const std::vector<int8_t> &kx0, const std::vector<int8_t> &ky0,
            const std::vector<int> &nbx0, const std::vector<int> &nby0, const std::vector<double> &kkx0,
            const std::vector<double> &kky0);
  // restrict fields from the finest level (FP64 node arrays) and refresh
  // diagonals and the coarse inverse
  int refresh(const double *nu, const double *nd, const double *ro, const double *gp, double mu, double dt,
              hipStream_t s, hipStream_t side = nullptr);
  // with a side stream, refresh leaves the coarsest inversion running there
  // (and refresh_from of an FP32 copy queues its inverse copy behind it);
  // join(s) makes s wait for both before the first V-cycle
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int join(hipStream_t s);
  bool tail_at(int l) const;
  // first level of the cooperative coarse V-cycle (-1: none): FP32, every
  // level from it to the tail a 5x5 Galerkin level of at most 32 default
  // tiles, at most VC_MAXL levels, all blocks co-resident
  int coop_level() const;
  int vcoarse(int l, hipStream_t s);
  DBuf<char> vcd;           // its arguments on the device, their pinned host copy
  char *vch = nullptr;      // and the event of its last upload
  hipEvent_t vce = nullptr;
  long vgen = 0, vk_gen = -1;  // hierarchy generation (set-up, refresh), and that of the upload
  int vk_l = -1, vk_ns = -1;
  double vk_om = 0;
  // the same hierarchy rounded from an FP64 one on the same geometry: the
  // Galerkin products and the coarsest inverse are formed in FP64 (products
  // of rounded fine-level entries lose the near-cancellation that carries
  // rigid-body modes of stiff regions), only the stored operators are T
  int refresh_from(const VelMG<double> &d, hipStream_t s);
  // z = M^-1 r for the velocity block; f/u live in L[0].fx/fy and L[0].ux/uy
  int vcycle(int l, hipStream_t s);
  // Tile shape of each Galerkin level's smoothing pass (candidate index,
  // 0 = the default): tune() times every candidate on the level's current
  // data (reps launches each, best kept) unless a level of that size was
  // tuned before in this process
  std::vector<int> shape;
  int tune(hipStream_t s, int reps = 10);
  VLev<T> lev(int l) const;
  void release();
};

// ---------------------------------------------------------------- scalar MG
template <class T>
struct ScalLevel {
  int X, Y, N;
  DBuf<double> cx, cy;   // point coordinates (device)
  std::vector<double> hcx, hcy;
  DBuf<T> k, c, dm;
  DBuf<int8_t> mask;
  std::vector<int8_t> hmask;
  DBuf<T> u, f, r, u2;  // u2: output of the post-smoothing pass (swapped with u)
  Map1 mx, my;
};

template <class T>
struct ScalMG {
  std::vector<ScalLevel<T>> L;
  int nc = 0;
  DBuf<double> dense;
  DBuf<double> ainv;  // FP64 in every precision
  std::vector<DBuf<int>> sfi, sfj;
  int nsmooth = 2;
  double omega = 1.0;       // damping of the smoothers' colour steps (exact rows undamped)
  // coords of the finest points; fi/fj give the coarse point selection
  int setup(const std::vector<double> &x0, const std::vector<double> &y0, int coarse_max);
  // finest-level fields: k, c, mask, dm (FP64 device arrays, length N0)
  int refresh(const double *k, const double *c, const int8_t *mask, const double *dm, hipStream_t s);
  int vcycle(int l, hipStream_t s);
  POp<T> op(int l) const;
  void release();
};

// ---------------------------------------------------------------- cell MG
// Galerkin multigrid for a scalar 5-point operator on the pressure cells
// (the BFBt operator B C^-1 B^T of the Stokes Schur preconditioner): 3x3
// stencils (slot-major, 9 slots), rows with mask = 1 are identity rows,
// tiled colour-Jacobi smoothing, hierarchy formed in FP64 and rounded to T.
template <class T>
struct CellLevel {
  int X, Y, N;
  DBuf<T> S;           // 9 x N
  DBuf<int8_t> mask;   // 1: identity row
  DBuf<T> u, u2, f, r;
  Map1 mx, my;         // to the next coarser level
  bool fused = false;  // residual and restriction in one kernel (tile spans fit its LDS tile)
};

template <class T>
struct CellMG {
  std::vector<CellLevel<T>> L;
  int nc = 0;
  DBuf<double> dense, ainv;
  DBuf<int> overflow;
  int nsmooth = 2;
  int nsmooth_c = 0;        // sweeps on the coarser levels (0: nsmooth)
  int tail_l = -1;          // first level run by the single-workgroup tail kernel (-1: none)
  DBuf<uint4> timg;         // the tail's LDS image (coarse inverse, stencils, row kinds, maps)
  int setup(const std::vector<double> &x0, const std::vector<double> &y0, int coarse_max);
  // copies the current coarse inverse, stencils and row kinds into timg
  int tail_pack(hipStream_t s);
  // fine stencil (FP64, 9 x N0) and mask given on device; forms the
  // Galerkin hierarchy and the coarsest inverse in FP64
  int refresh(const double *S0, const int8_t *mask0, hipStream_t s);
  // the same hierarchy rounded from an FP64 one
  int refresh_from(const CellMG<double> &d, hipStream_t s);
  // one V-cycle: u = M^-1 f on L[0].f -> L[0].u
  int vcycle(int l, hipStream_t s);
  void release();
};

// coarse line selection: every second line plus the last
std::vector<int> coarse_lines(int n);
int build_geometry(const double *gx, const double *gy, int X, int Y, int coarse_max, std::vector<GeoLevel> &geo);

}  // namespace i2e
