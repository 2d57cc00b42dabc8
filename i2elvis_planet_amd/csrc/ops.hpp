// ops.hpp -- FP64 node kernels (exact operators, right-hand sides, reductions)
#pragma once
#include "kernels.hpp"
#include "rows.hpp"

namespace i2e {

// continuity-type velocity BC (movemart.c:863-879): d = (first neighbour's
// unknown) - (this unknown), 0-based; the row is the continuity equation of
// node (m1 + dm1, m2 + dm2).  C integer division, as the reference's.
__host__ __device__ inline void cont_shift(int64_t d, int64_t Y, int &dm1, int &dm2) {
  const int64_t m3 = d / 3;
  dm1 = m3 >= Y;
  dm2 = m3 == 1 || m3 > Y;
}


struct GravGeo {
  double R, gxkoef, mu;
};
// distance of cell (i, j)'s centre from the box centre (movemart.c:636-650)
__device__ __forceinline__ double grav_dist(const Grid &g, int i, int j) {
  double dx = (g.gx[i - 1] + g.gx[i] - g.xsize) / 2.0, dy = (g.gy[j - 1] + g.gy[j] - g.ysize) / 2.0;
  return pow(dx * dx + dy * dy, 0.5);
}

__global__ void k_stokes_apply(SOp<double> o, BcTab bc, const double *x, double *y);
__global__ void k_stokes_rhs(SOp<double> o, BcTab bc, double *b);
__global__ void k_stokes_pack(int X, int Y, const double *pr, const double *vx, const double *vy, double *x);
__global__ void k_stokes_unpack(int X, int Y, const double *x, double *pr, double *vx, double *vy);
__global__ void k_strain(Grid g, NodeF F);
__global__ void k_maxvel(int X, int Y, const double *vx, const double *vy, double *out);
__global__ void k_grav_mean(Grid g, const double *ro, double R, double *part);
__global__ void k_grav_apply(Grid g, GravGeo q, const double *x, double *y);
__global__ void k_grav_rhs(Grid g, GravGeo q, const double *ro, double *b);
__global__ void k_grav_rowscale(Grid g, GravGeo q, double *sc);
__global__ void k_heat_rowscale(Grid g, NodeF F, BcTab bc, double dt, double *sc);
__global__ void k_grav_ga(Grid g, const double *gp, double *ga);
__global__ void k_heat_source(const i2e_params *p, Grid g, NodeF F, BcTab bc, double timesum, double *sol1,
                              double *out);
__global__ void k_heat_apply(Grid g, NodeF F, BcTab bc, double dt, const double *x, double *y);
__global__ void k_heat_rhs(Grid g, NodeF F, BcTab bc, double dt, const double *sol1, double *rhs);

}  // namespace i2e
