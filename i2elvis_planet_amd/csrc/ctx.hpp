// ctx.hpp -- the context behind the opaque i2e_ctx handle of include/i2e.h,
// shared by the translation units that implement its entry points.
#pragma once
#include "krylov.hpp"

namespace i2e {
struct AdjHist;               // recorded trajectory and thermal history (adjoint.hip)
void adj_release(AdjHist *h);
struct BandLU;                // banded LU of the Stokes rows (band.hip)
void band_release(BandLU *b);
__global__ void k_copy(int n, const double *a, double *b);
__global__ void k_heat_fields(int N, const double *ro, const double *cp, double dt, const int *bondm, double *c,
                              int8_t *mask, double *dm);
// deterministic two-sum reduction of adjoint.hip: k_sum_part over RG blocks
// of RB threads, then k_sum_final (fixed order)
constexpr int RB = 256, RG = 1024;
__global__ void k_sum_part(int64_t n, const double *a, const double *b, double *part);
__global__ void k_sum_final(const double *part, double *out);
}  // namespace i2e

using namespace i2e;

struct i2e_ctx {
  i2e_params P;
  i2e_params *dP = nullptr;
  int X = 0, Y = 0, N = 0;
  int64_t Mmax = 0, M = 0;
  hipStream_t s = nullptr;
  hipStream_t s2 = nullptr;  // second stream: the velocity coarsest inversion overlaps the Stokes set-up
  DBuf<double> gx, gy, nodes, acc, work, sol1, tk2, red, xs, bs;
  DBuf<float> mfl;
  DBuf<int8_t> mt;
  DBuf<int> bondm, bondn;
  DBuf<double> bondv;
  NodeF F;
  Markers mk;
  Acc A;
  BcTab bc;
  int need_eii = 0;
  bool setup_done = false;
  std::vector<GeoLevel> geo;
  VelMG<float> v32;
  VelMG<double> v64;
  CellMG<float> l32;
  CellMG<double> l64;
  DBuf<double> bcix, bciy, bLS;  // BFBt: C^-1 on momentum rows, FP64 cell stencil
  DBuf<int8_t> blmask;
  ScalMG<float> h32, g32;
  ScalMG<double> h64, g64;
  DBuf<double> hk, hc, hdm, gk, gc, gdm, gsc, hsc;
  DBuf<int8_t> hmask, gmask;
  Fgmres fs, fh, fgv;
  i2e_solver_opts o;
  i2e_stats st;
  double timestep = 0, timesum = 0, gxkoef = 1.0;
  HBuf<double> pin;
  KTimer kt;
  AdjHist *adj = nullptr;       // allocated by i2e_record
  int heat_solved = 0;          // the last i2e_titerate ran its solve and marker update
  DBuf<double> xprev;           // Stokes solution of the step before the last (extrapolated guess)
  int have_prev = 0;            // xprev holds a solution of this run (reset by i2e_upload)
  double dt_prev = 0;           // timestep of the last Stokes solve
  DBuf<int> tcnt;               // per-cell marker-type counts (i2e_type_counts)
  DBuf<unsigned long long> hacc;  // body-statistics counters (i2e_hydration)
  DBuf<double> hdacc;           // its sums and the largest node temperature
  // cell sort of the markers (sort.hip): second marker storage for the
  // permutation, the caller's index of each slot, keys, counts/offsets
  DBuf<float> mfl2;
  DBuf<int8_t> mt2;
  DBuf<int> morig, morig2, mkey, mcnt, moff, mbsum;
  DBuf<int> mslot0;             // slot of the caller's marker 0 (while permuted)
  // magnetisation per marker slot (markmg_old, markmg_time) and their spares
  DBuf<int8_t> mg, mg2;
  DBuf<double> mgt, mgt2;
  int permuted = 0;             // slots are not in the caller's order (morig maps them back)
  int tbc_passes = 1;            // tkrecalc projection passes (longest chain of temperature BC rows)
  int sort_every = 10;          // i2e_step sorts the markers every this many steps (0: never)
  int since_sort = 0;
  // the node velocities, pressure and viscosities are the last Stokes
  // solve's, with the markers that set them unchanged since (the flow
  // adjoint differentiates that solve); stokes_dt is its stabilisation dt
  int stokes_valid = 0;
  double stokes_dt = 0;
  DBuf<double> fadj;            // flow adjoint scratch (flow_adjoint.hip)
  BandLU *band = nullptr;       // stokes_direct: the factored Stokes rows (band.hip)
  DBuf<int> fflag;
  DBuf<int8_t> fmask;

  Grid grid() const { return Grid{X, Y, P.xsize, P.ysize, gx.p, gy.p}; }
  SOp<double> sop() const { return SOp<double>{X, Y, gx.p, gy.p, F.nu, F.nd, F.ro, F.gp, timestep, P.nubeg * P.strmin}; }
};

// the Stokes preconditioner of the current solve (ctx.hip); presplit: the
// velocity part of r is already in the finest level's right-hand side
int stokes_precond(i2e_ctx *c, const double *r, double *z, bool presplit = false);
// the Stokes set-up of i2e_viterate for the context's current fields and
// timestep: right-hand side, initial guess, multigrid hierarchies (ctx.hip)
int stokes_prepare(i2e_ctx *c);
// G x = b for the gravity rows, from x = 0 (b overwritten; ctx.hip)
int gravity_solve_rhs(i2e_ctx *c, double *b, double *x, int *its, double *rr);
// stokes_direct: assemble, equilibrate and factor the Stokes rows of the
// current fields; z = A^-1 r by the factors (band.hip)
int band_factor(i2e_ctx *c);
int band_apply(i2e_ctx *c, const double *r, double *z, bool scaled = false);
// stokes_direct = 2: FGMRES on the equilibrated rows R A C y = R b
int band_scale_system(i2e_ctx *c, double *b, double *x);
int band_unscale(i2e_ctx *c, double *x);
int band_op(i2e_ctx *c, const double *y, double *out);
int band_dims(const i2e_ctx *c, int *kl, int *ku, double *bytes);
