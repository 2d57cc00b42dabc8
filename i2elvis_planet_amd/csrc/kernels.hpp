// kernels.hpp -- device data views and kernel declarations shared by the
// translation units of libi2e.
#pragma once
#include "common.hpp"

namespace i2e {

// node fields (reference names, headmart.c:87-99)
// node fields; the context allocates them as one block, N apart in this
// order (k_ronu_markers relies on it)
struct NodeF {
  double *vx, *vy, *pr, *gp, *ga;
  double *exx, *eyy, *exy, *sxx, *syy, *sxy;
  double *nu, *nd, *ro, *tk, *cp, *kt, *ht, *et;
};

// marker SoA (headmart.c:122-127); the float fields are one block, one
// capacity apart in this order (k_ronu_markers relies on it)
struct Markers {
  float *x, *y, *k, *v, *p, *por, *tmax;
  int8_t *t;
  int8_t *mg;   // magnetisation state (markmg_old)
  double *mgt;  // its time, Ma (markmg_time)
};

// ronurecalc accumulators (ro0..tk0, sol0/sol1 and the viscosity weights),
// one block, N apart in this order (k_ronu_markers relies on it)
struct Acc {
  double *ro0, *et0, *cp0, *kt0, *ht0, *tk0, *nu0, *nd0, *w0, *w1, *wnu, *wnd;
};

// boundary tables on device: bondm[4N] (row, 0 = none), bondn[3*nb]
// (1-based unknown numbers, 0 = none), bondv[4*nb]
struct BcTab {
  const int *bondm;
  const int *bondn;
  const double *bondv;
};

// radiogenic heating per kg of a step (markmart.c:758-785): silicate (26Al
// and, with longlived, 40K/235U/238U/232Th) and iron (60Fe, 40K)
struct HeatK {
  double sil, fe;
};
HeatK heat_per_kg(const i2e_params &p, double timesum);

// markers.hip
__global__ void k_movemark(const i2e_params *p, Grid g, NodeF F, Markers mk, int64_t M, double timestep, int need_eii);
// TILED: one block per marker tile of the last cell sort (toff: the sort's
// key offsets, ntj tiles per column of tiles), node sums staged in LDS
template <bool TILED, bool IMM>
__global__ void k_ronu_markers(const i2e_params *p, Grid g, NodeF F, Markers mk, int64_t M, Acc A, HeatK hk,
                               double timestep, double timesum, int need_eii, const double *tref, const int *toff,
                               int ntj);
__global__ void k_tref(const i2e_params *p, Grid g, Markers mk, const int *slot0, double *tref);
__global__ void k_ronu_nodes(const i2e_params *p, NodeF F, Acc A, int N, double *mm, int imm);
__global__ void k_contrast(const i2e_params *p, NodeF F, int N, const double *mm);
__global__ void k_tbc(NodeF F, BcTab bc, int N);
template <bool TILED>
__global__ void k_heat_markers_a(const i2e_params *p, Grid g, NodeF F, const double *tk2, Markers mk, int64_t M,
                                 double timestep, double *sol0, double *sol1, const int *toff, int ntj);
__global__ void k_heat_nodes_norm(int N, const double *sol0, double *sol1);
__global__ void k_heat_markers_b(Grid g, Markers mk, int64_t M, const double *sol1);
__global__ void k_type_counts(Grid g, Markers mk, int64_t M, int ntypes, int *cnt);
// hydration and body statistics (impact.c:3022-3228): acc[0..16) marker
// counters, acc[16] bits of the largest silicate marker T above the ambient,
// acc[17] body node count; dacc[0] sum of silicate marker T, dacc[1] sum of
// body node T, dacc[2] largest node T
__global__ void k_hyd_nodes(int N, const double *tk, double tamb, unsigned long long *acc, double *dacc);
__global__ void k_hyd_markers(Markers mk, int64_t M, double t_sol, double t_liq, int hydrate, const double *maxtk,
                              double tamb, unsigned long long *acc, double *dacc);

// sort.hip (exclusive scans run in blocks of SB elements: 256 threads x 4)
constexpr int SB = 1024;
// The sort key is tile-major: kTile x kTile cells form a tile, tiles are
// numbered column-major (ntj per column), cells inside a tile column-major.
// A tile's markers are then one contiguous slot range, and the scatter of
// one tile touches (kTile+1)^2 nodes.
constexpr int kTile = 8;
__host__ __device__ __forceinline__ int tiles_per_col(int Y) { return (Y - 1 + kTile - 1) / kTile; }
__host__ __device__ __forceinline__ int tile_count(int X, int Y) {
  return ((X - 1 + kTile - 1) / kTile) * tiles_per_col(Y);
}
__device__ __forceinline__ int tile_key(int ci, int cj, int ntj) {
  return (((ci / kTile) * ntj + cj / kTile) * kTile + ci % kTile) * kTile + cj % kTile;
}
__global__ void k_mkey(Grid g, Markers mk, int64_t M, int C, int *key, int *cnt);
__global__ void k_scan_blocks(int n, const int *in, int *out, int *bsum);
__global__ void k_scan_sums(int nb, int *bsum);
__global__ void k_scan_add(int n, const int *bsum, int *out);
__global__ void k_mscatter(Markers src, Markers dst, int64_t M, const int *key, int *off, const int *orig,
                           int *orig2, int *slot0);
__global__ void k_munperm(Markers src, Markers dst, int64_t M, const int *orig);

}  // namespace i2e
