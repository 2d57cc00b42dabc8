// markers.hip -- per-marker kernels of the step: RK4 advection (movemark,
// markmart.c:3-188), marker properties + marker->node scatter and node
// normalisation (ronurecalc, markmart.c:192-1009), temperature boundary
// projection (tkrecalc, heatmart.c:758-826) and the marker temperature update
// after the implicit heat solve (titerate, heatmart.c:515-689).
//
// Markers are kept grouped by cell and tile (the cell sort, sort.hip) and
// move at most maxxystep cells per step, so a wavefront touches a handful of
// neighbouring cells: node reads hit L2, and the scatters reduce same-cell
// lanes in registers before an FP64 atomic add (to LDS for the tiled
// ronurecalc scatter, to HBM otherwise).
#include "kernels.hpp"
#include "viscalc.hpp"

// Waves per SIMD the ronurecalc marker kernel is compiled for: 3 caps it at
// 168 VGPRs (8 spilled) and measured 12-15 % faster than 2 (211 VGPRs) at
// 6.25M markers; 4 spills 54 and is slower again.
#ifndef I2E_RONU_WAVES
#define I2E_RONU_WAVES 3
#endif

namespace i2e {

// exp(kp P), the pressure factor of the conductivity (markmart.c:486):
// exactly 1 when kp = 0 and P is finite, without the exponential
__device__ __forceinline__ double exp_kp(double kp, double mpb) {
  return (kp == 0 && isfinite(mpb)) ? 1.0 : exp(kp * mpb);
}

// allintere, markmart.c:1458-1543: strain rate, stress, P and |g| at (x,y)
struct Ev {
  double exy, sxy, exx, sxx, eyy, syy, pr, ga;
  int m10, m20;
};

__device__ __forceinline__ void clampxy(const Grid &g, double &x, double &y) {
  if (x < 0)
    x = 0;
  else if (x > g.xsize)
    x = g.xsize;
  if (y < 0)
    y = 0;
  else if (y > g.ysize)
    y = g.ysize;
}

// (|g| at the marker is only read by the soft-turbulence conductivity:
// with_ga = false skips that gather)
__device__ void allintere_d(const Grid &g, const NodeF &F, double x, double y, Ev &e, bool full,
                            bool with_ga = false) {
  W4 w;
  clampxy(g, x, y);
  e.m10 = serch(g.gx, g.X, x);
  e.m20 = serch(g.gy, g.Y, y);
  if (full) {
    nodewt2(g.gx, g.gy, bn_i0(e.m10, g.X), bn_i0(e.m20, g.Y), x, y, 0, 0, w);
    e.exy = isum(w, g.Y, F.exy);
    e.sxy = isum(w, g.Y, F.sxy);
  }
  nodewt2(g.gx, g.gy, ce_i0(g.gx, e.m10, x, g.X), ce_i0(g.gy, e.m20, y, g.Y), x, y, -1, -1, w);
  if (full) {
    e.exx = isum(w, g.Y, F.exx);
    e.sxx = isum(w, g.Y, F.sxx);
    e.eyy = isum(w, g.Y, F.eyy);
    e.syy = isum(w, g.Y, F.syy);
  }
  e.pr = isum(w, g.Y, F.pr);
  e.ga = with_ga ? isum(w, g.Y, F.ga) : 0.0;
}

// allinterv, markmart.c:1378-1453
__device__ __forceinline__ void allinterv_d(const Grid &g, const double *vx, const double *vy, double x, double y,
                                            double &ovx, double &ovy) {
  W4 w;
  int m10, m20;
  clampxy(g, x, y);
  m10 = serch(g.gx, g.X, x);
  m20 = serch(g.gy, g.Y, y);
  nodewt2(g.gx, g.gy, m10, vx_j0(g.gy, g.Y, m20, y), x, y, 0, +1, w);
  ovx = isum(w, g.Y, vx);
  nodewt2(g.gx, g.gy, vy_i0(g.gx, g.X, m10, x), m20, x, y, +1, 0, w);
  ovy = isum(w, g.Y, vy);
}

// ---------------------------------------------------------------- movemark
__global__ __launch_bounds__(256) void k_movemark(const i2e_params *__restrict__ p, Grid g, NodeF F, Markers mk,
                                                 int64_t M, double timestep, int need_eii) {
  for (int64_t mm1 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; mm1 < M; mm1 += (int64_t)gridDim.x * blockDim.x) {
    int t = mk.t[mm1];
    int mm2 = t >= 100 ? t - 100 : t;
    float fx = mk.x[mm1], fy = mk.y[mm1];
    if (!((fx >= 0 && fy >= 0 && (double)fx <= g.xsize && (double)fy <= g.ysize) || p->outgrid != 1)) continue;
    if ((unsigned)mm2 < (unsigned)I2E_MAXTYPES && p->immobile[mm2]) continue;  // markmart.c:27
    Ev e;
    allintere_d(g, F, (double)fx, (double)fy, e, need_eii != 0);
    double mpb = e.pr * 1e-5, mtk = (double)mk.k[mm1];
    mk.p[mm1] = (float)e.pr;
    if (t < 50 && mtk > 0) {
      double eii = need_eii ? pow(0.5 * (e.exx * e.exx + e.eyy * e.eyy) + e.exy * e.exy, 0.5) : 0.0;
      mk.v[mm1] = (float)viscalc_d(p, mtk, mpb, t < 20 ? mm2 : mm2 - 20, eii);
    }
    double x0 = (double)fx, y0 = (double)fy, vx1, vy1, vx2, vy2, vx3, vy3, vx4, vy4;
    allinterv_d(g, F.vx, F.vy, x0, y0, vx1, vy1);
    double vx0 = vx1, vy0 = vy1;  // markmod 1: Euler (markmart.c:74-79)
    if (p->markmod != 1) {        // otherwise RK4 (markmart.c:80-97)
      allinterv_d(g, F.vx, F.vy, x0 + vx1 * timestep / 2.0, y0 + vy1 * timestep / 2.0, vx2, vy2);
      allinterv_d(g, F.vx, F.vy, x0 + vx2 * timestep / 2.0, y0 + vy2 * timestep / 2.0, vx3, vy3);
      allinterv_d(g, F.vx, F.vy, x0 + vx3 * timestep, y0 + vy3 * timestep, vx4, vy4);
      vx0 = (vx1 + 2.0 * vx2 + 2.0 * vx3 + vx4) / 6.0;
      vy0 = (vy1 + 2.0 * vy2 + 2.0 * vy3 + vy4) / 6.0;
    }
    if (p->outgrid == 2) {
      if (fx < 0 || (double)fx > g.xsize) vy0 = 0;
      if (fy < 0 || (double)fy > g.ysize) vx0 = 0;
    }
    fx += (float)(timestep * vx0);
    fy += (float)(timestep * vy0);
    mk.x[mm1] = fx;
    mk.y[mm1] = fy;
    if (fx < 0 || fy < 0 || (double)fx > g.xsize || (double)fy > g.ysize) mk.k[mm1] = 0;
  }
}

// ---------------------------------------------------------------- ronurecalc
HeatK heat_per_kg(const i2e_params &p, double timesum) {
  HeatK h{0.0, 0.0};
  const double timesumA = timesum / (365.250 * 24.000 * 3600.000);
  const double init_al = 1.89716e+23 * p.al2627_init * 4.999e-13 / (3.264e+13);
  double ll = 0;
  if (p.longlived != 0)
    ll = 0.40689 * 1.430e-11 * exp(-timesumA * 4.9867e-10) + 2.990e-12 * exp(-timesumA * 9.8458e-10) +
         1.600e-12 * exp(-timesumA * 1.5541e-10) + 1.000e-12 * exp(-timesumA * 4.9511e-11);
  h.sil = init_al * exp(-timesumA * 9.6673e-7) * p.h0_scale + ll;
  const double init_fe = 1.971e+24 * p.fe6056_init * 4.3451e-13 / (1.183e+14);
  double ka = 0;
  if (p.longlived != 0) ka = 0.090625 * 1.430e-11 * exp(-timesumA * 4.9867e-10);
  h.fe = init_fe * exp(-timesumA * 2.6660e-7) + ka;
  return h;
}

// Henke et al. cold pressing / sintering (markmart.c:363-468); the Newton
// restarts the reference repeats without bound stop after 64 here.
__device__ void porosity_update_d(const i2e_params *p, float &por, float markk, double mpb, double mtk,
                                  double timestep) {
  const double E_act = 85000.000, fact_a = 4.000e-5, grain_orig = 1.000e-6, pivalue = 3.141592654, r_gas = 1.9872,
               p_ref = 0.130;
  if (markk < 700.000) {
    double lim = 0.420 + 0.460 * pow((pow((mpb / p_ref), 1.720) + 1.000), (-1.000));
    double pd = (double)por;
    por = (float)(pd <= lim ? pd : lim);
    pd = (double)por;
    por = (float)(pd <= 0.850 ? pd : 0.850);
  }
  if (markk >= 700.000) {
    double grain_corr = 0, grain_guess, grain_old = 0, por_change = 0, por_orig, funct_error;
    int n_count = 1;
    por_orig = 0.420 + 0.460 * pow((pow((mpb / p_ref), 1.720) + 1.000), (-1.000));
    if (p->por_init <= por_orig) por_orig = p->por_init;
    if (por_orig > 0.850) por_orig = 0.850;
    grain_guess = grain_orig;
    double q = (1.000 - por_orig) / (1.000 - (double)por);
    for (;;) {
      for (int it = 1; it < 100; it++) {
        grain_old = grain_guess;
        double funct = pow(grain_guess, 3.000) + pow(grain_orig, 3.000) -
                       3.000 * pow(grain_guess, 2.000) * grain_orig * pow(q, (1.000 / 3.000)) +
                       pow(grain_orig, 3.000) * (1.000 - por_orig) / (1.000 - (double)por);
        double funct_drv = 3.000 * pow(grain_guess, 2.000) - 6.000 * grain_guess * grain_orig * pow(q, (1.000 / 3.000));
        grain_guess = grain_old - (funct / funct_drv);
      }
      funct_error = dabs(grain_guess - grain_old);
      if (((funct_error > 5.000e-2) || (grain_guess <= grain_orig)) && n_count < 64) {
        n_count++;
        grain_guess = 1.250 * (double)n_count * grain_orig;
        continue;
      }
      break;
    }
    if (grain_guess > grain_orig) grain_corr = grain_guess;
    double cont_r = sqrt(pow(grain_corr, 2.000) - pow(grain_orig, 2.000) * pow(q, (2.000 / 3.000)));
    double c_fact = 2.000 * sqrt(3.000) * (pow(grain_corr, 2.000) - pow(cont_r, 2.000));
    double grain_stress = c_fact * mpb / (pivalue * pow(cont_r, 2.000));
    if (por > 0.100)
      por_change = fact_a * (1.000 - (double)por) * pow(grain_stress, 1.500) / pow((100.000 * grain_corr), 3.000) *
                   exp(-E_act / (r_gas * mtk)) * timestep;
    if (por <= 0.100)
      por_change = 10.000 * (double)por * fact_a * (1.000 - (double)por) * pow(grain_stress, 1.500) /
                   pow((100.000 * grain_corr), 3.000) * exp(-E_act / (r_gas * mtk)) * timestep;
    if (por_change > 0.000) por = (float)((double)por - por_change);
  }
}

// Scatter contributions of one marker: per corner (a,b) of its cell the ten
// weighted sums of markmart.c:840-911, plus the cell-centre weights of nd.

// soft-turbulence conductivity (markmart.c:503-610): the heat flux of soft
// turbulence, q = 0.089 k^(2/3) |dT|^(4/3) (rho^2 beta g cp / eta)^(1/3),
// with the fluid's viscosity eta_f, and the conductivity that gives the same
// q with the numerical viscosity mnu; kt itself when q is tiny.  dT is the
// float difference of the marker temperature and marker 0's.
__device__ double soft_turb_k(double kt, float dTf, double mro, double bb, double mga, double mcp, double eta_f,
                              double mnu) {
  const double heat_flux = 0.089 * pow(kt, (2.000 / 3.000)) * pow((double)(dTf >= 0 ? dTf : -dTf), (4.000 / 3.000)) *
                           pow((mro * mro * bb * mga * mcp / eta_f), (1.000 / 3.000));
  if (heat_flux <= 0.0001) return kt;
  return pow((heat_flux / 0.089), (3.000 / 2.000)) / (pow((double)dTf, 2.000) * mro) *
         pow((bb * mga * mcp / mnu), (-1.000 / 2.000));
}

// the temperature of the caller's marker 0 as the marker loop of ronurecalc
// sees it after that marker's own turn (sticky air reset to tmp_ambient,
// markmart.c:296-309): the reference temperature of the soft-turbulence dT
__global__ void k_tref(const i2e_params *__restrict__ p, Grid g, Markers mk, const int *slot0, double *tref) {
  if (blockIdx.x || threadIdx.x) return;
  const int i = slot0 ? *slot0 : 0;
  const int t = mk.t[i];
  const int mm2 = t >= 100 ? t - 100 : t;
  float fk = mk.k[i];
  const float fx = mk.x[i], fy = mk.y[i];
  if (fx > 0 && fy > 0 && (double)fx < g.xsize && (double)fy < g.ysize && fk > 0 && t < 50 && mm2 < 2)
    fk = (float)p->tmp_ambient;
  *tref = (double)fk;
}

// Node sums per run of same-cell lanes (segmented sums over the wavefront),
// added into the twelve accumulators: straight to HBM with atomics
// (untiled), or -- TILED, one block per marker tile of the last sort -- into
// an LDS copy of the tile's (kTile+1)^2 nodes, flushed with one atomic per
// node and field at the end.  Markers that left their tile since the sort
// take the atomics to HBM; any order stays correct.
// IMM: some rock type is immobile: its markers' temperatures stay out of the
// node temperature, which then has a weight sum of its own (w1)
template <bool TILED, bool IMM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(I2E_RONU_WAVES))) void k_ronu_markers(const i2e_params *__restrict__ p, Grid g, NodeF F, Markers mk,
                                                     int64_t M, Acc A, HeatK hk, double timestep, double timesum,
                                                     int need_eii, const double *__restrict__ tref,
                                                     const int *__restrict__ toff, int ntj) {
  constexpr int TN = kTile + 1, TA = TN * TN, NF = IMM ? 12 : 11;
  // LDS node sums of the tile: nu0 wnu ro0 et0 cp0 kt0 ht0 w0 tk0 nd0 wnd (w1)
  __shared__ double tacc[TILED ? NF * TA : 1];
  const int lane = threadIdx.x & 63;
  int ti0 = 0, tj0 = 0;
  int64_t lo, hi, stride;
  if (TILED) {
    const int kk = kTile * kTile, T = blockIdx.x;
    lo = T ? toff[T * kk - 1] : 0;
    hi = toff[T * kk + kk - 1];
    stride = blockDim.x;
    ti0 = (T / ntj) * kTile;
    tj0 = (T % ntj) * kTile;
    for (int k = threadIdx.x; k < NF * TA; k += blockDim.x) tacc[k] = 0;
    __syncthreads();
  } else {
    lo = (int64_t)blockIdx.x * blockDim.x;
    hi = M;
    stride = (int64_t)gridDim.x * blockDim.x;
  }
  {
    // Every field pointer the kernel uses, derived from the three block
    // bases the context allocates (node fields and accumulators N apart in
    // NodeF / Acc order, marker float fields one capacity apart): passed as
    // separate kernel arguments they outgrow the scalar registers and spill
    // into vector-register lanes, one readlane per use.
    const size_t N = (size_t)g.X * g.Y, MS = mk.y - mk.x;
    double *const n0 = F.vx, *const a0 = A.ro0;
    F.pr = n0 + 2 * N;
    F.ga = n0 + 4 * N;
    F.exx = n0 + 5 * N;
    F.eyy = n0 + 6 * N;
    F.exy = n0 + 7 * N;
    F.sxx = n0 + 8 * N;
    F.syy = n0 + 9 * N;
    F.sxy = n0 + 10 * N;
    A.et0 = a0 + N;
    A.cp0 = a0 + 2 * N;
    A.kt0 = a0 + 3 * N;
    A.ht0 = a0 + 4 * N;
    A.tk0 = a0 + 5 * N;
    A.nu0 = a0 + 6 * N;
    A.nd0 = a0 + 7 * N;
    A.w0 = a0 + 8 * N;
    A.w1 = a0 + 9 * N;
    A.wnu = a0 + 10 * N;
    A.wnd = a0 + 11 * N;
    mk.k = mk.x + 2 * MS;
    mk.v = mk.x + 3 * MS;
    mk.p = mk.x + 4 * MS;
    mk.por = mk.x + 5 * MS;
    mk.tmax = mk.x + 6 * MS;
  }
  // whole wavefronts iterate together: the scatter reduces across lanes
  for (int64_t base = lo; base < hi; base += stride) {
    const int64_t mm1 = base + threadIdx.x;
    bool act = mm1 < hi;
    // the marker's properties and its four scatter weights (0 off the
    // active markers); the per-corner products are formed at the scatter
    double s_ro = 0, s_bb = 0, s_cp = 0, s_kt = 0, s_ht = 0, s_tk = 0, s_nu = 0, s_mob = 0;
    double wx0 = 0, wx1 = 0, wy0 = 0, wy1 = 0, c_nd = 0, c_wnd = 0;
    int key = -1;
    if (act) {
      int t = mk.t[mm1];
      int mm2 = t >= 100 ? t - 100 : t;
      float fx = mk.x[mm1], fy = mk.y[mm1], fk = mk.k[mm1];
      if (!(fx > 0 && fy > 0 && (double)fx < g.xsize && (double)fy < g.ysize && fk > 0 && t < 50)) act = false;
      if (act) {
        const double x = (double)fx, y = (double)fy;
        Ev e;
        allintere_d(g, F, x, y, e, need_eii != 0, p->soft_turb != 0);
        double mpb = e.pr * 1.0000e-5, mtk = (double)fk;
        if (mm2 < 2) {
          fk = (float)p->tmp_ambient;
          mk.k[mm1] = fk;
          mtk = (double)fk;
        }
        double mnu = mk.v[mm1];
        if (mnu <= 0) {
          double eii = need_eii ? pow(0.5 * (e.exx * e.exx + e.eyy * e.eyy) + e.exy * e.exy, 0.5) : 0.0;
          mnu = viscalc_d(p, mtk, mpb, t < 20 ? mm2 : mm2 - 20, eii);
        }
        double celdx = fx - g.xsize / 2.0, celdy = fy - g.ysize / 2.0;
        celdy = sqrt(celdx * celdx + celdy * celdy);
        if (celdy > g.xsize / 2.0 * p->radmax_frac) mnu = p->eta_outer;
        float por = mk.por[mm1];
        if (mm2 == 0 || mm2 >= 7) por = 0.000f;
        if (t != 0)
          if (mk.tmax[mm1] < fk) mk.tmax[mm1] = fk;
        if (p->porosity_on && p->porous[mm2] && (por > 0.000f && por <= 0.850f))
          porosity_update_d(p, por, fk, mpb, mtk, timestep);
        if (por < 0.000f) por = 0.000f;
        if (por > 1.000f) por = 1.000f;
        mk.por[mm1] = por;
        const i2e_rock &r = p->rock[mm2];
        double mro, mbb = 0;
        if (p->densimod == 0)
          mro = r.ro;
        else {
          mbb = r.bb / (1.0 - r.bb * (mtk - 298.15));
          mro = r.ro * (1.0 - r.bb * (mtk - 298.15)) * (1.0 + r.aa * (mpb - 1.0) * 1e-3) * (1.000 - (double)por);
        }
        double mcp = r.cp;
        double mkt = (r.kt + r.kf / (mtk + 77.0)) * exp_kp(r.kp, mpb);
        double xmelt = 0;  // the marker's own silicate melt fraction
        if (p->melt_on && p->meltable[mm2 < 20 ? mm2 : mm2 - 20]) {
          double xm = (mtk - p->melt_ts) / (p->melt_tl - p->melt_ts);
          if (xm < 0.0) xm = 0.0;
          if (xm > 1.0) xm = 1.0;
          xmelt = xm;
          int tnew = t;
          if (xm > 0 && tnew < 20) tnew += 20;
          if (xm <= 0 && tnew > 20) tnew -= 20;
          mk.t[mm1] = (int8_t)tnew;
          if (tnew > 20) {
            const i2e_rock &rm = p->rock[tnew];
            const i2e_rock &rs = p->rock[tnew - 20];
            mro = mro * (1.0 - xm) + xm * rm.ro * (1.0 - rm.bb * (mtk - 298.15)) * (1.0 + rm.aa * (mpb - 1.0) * 1e-3);
            if (xm > p->melt_switch) {
              double nueff = p->melt_eta;
              if (nueff < p->nubeg) nueff = p->nubeg;
              if (nueff > p->nuend) nueff = p->nuend;
              mnu = nueff;
            }
            mkt = ((rm.kt + rm.kf / (mtk + 77.0)) * exp_kp(rm.kp, mpb)) * xm +
                  ((rs.kt + rs.kf / (mtk + 77.0)) * exp_kp(rs.kp, mpb)) * (1.0 - xm);
            if (xm > 0 && xm < 1.0) {
              double xp = (mtk + 1.0 - p->melt_ts) / (p->melt_tl - p->melt_ts);
              double xq = (mtk - 1.0 - p->melt_ts) / (p->melt_tl - p->melt_ts);
              xp = xp < 0.0 ? 0.0 : (xp > 1.0 ? 1.0 : xp);
              xq = xq < 0.0 ? 0.0 : (xq > 1.0 ? 1.0 : xq);
              mcp += (p->melt_hlat * xp - p->melt_hlat * xq) / (2.0 * 1.0);
            }
          }
        }
        if (p->soft_turb && timesum > p->core_form_time) {
          const float dTf = fk - (float)*tref;
          if (mm2 == 25 || mm2 == 26) {
            // crystal mush between 40 and 60 % melt, melt above
            double eta_f = -1.0;
            if (xmelt >= 0.4000 && xmelt <= 0.6000) eta_f = r.n0 * exp((-172.690) * (xmelt - 0.400));
            if (xmelt > 0.6000) eta_f = 100.0000;
            if (eta_f > 0) mkt = soft_turb_k(r.kt, dTf, mro, r.bb, e.ga, mcp, eta_f, mnu);
          }
          if (p->is_iron[mm2]) {
            double xfe = 0;
            if (p->fe_tl > 0) {
              xfe = (mtk - p->fe_ts) / (p->fe_tl - p->fe_ts);
              if (xfe < 0.000) xfe = 0.0000;
              if (xfe > 1.000) xfe = 1.0000;
            }
            // solid iron 1e12 Pa s, molten iron 1e-2 Pa s
            mkt = soft_turb_k(r.kt, dTf, mro, r.bb, e.ga, mcp, xfe < 1.0000 ? 1.0000e12 : 0.010, mnu);
          }
        }
        if (p->porous[mm2]) {
          double kt0 = r.kt, pr0 = (double)por;
          if (por < 0.200f) mkt = kt0 * exp(-pr0 / 0.080);
          if (por >= 0.200f && por <= 0.400f)
            mkt = pow((pow((kt0 * exp(-pr0 / 0.080)), 4.000) + pow((kt0 * exp(-1.200 - (pr0 / 0.167))), 4.000)), 0.250);
          if (por > 0.400f) mkt = kt0 * exp(-1.200 - (pr0 / 0.167));
        }
        if (mkt <= 0.0010) mkt = 0.0010;
        if (mkt >= p->k_cutoff) mkt = p->k_cutoff;
        if (p->magnet) {
          // Curie temperatures of magnetite (851.16 K) and hematite (948.16 K);
          // states 0 none, 1 hematite, 2 both, 3 cannot magnetise
          const int old = mk.mg[mm1];
          const bool sil = mm2 == 5 || mm2 == 6;
          int mag = 3;
          double mag_time = mk.mgt[mm1];
          const double now = timesum / (3600.000 * 24.000 * 365.250 * 1.000e+6);
          if (p->dynamo == 0) {
            if (sil) {
              if (mtk <= 851.16 && old == 2)
                mag = 2;
              else if ((mtk > 851.16 && mtk <= 948.16) && old == 2)
                mag = 1;
              else if (mtk <= 851.16 && old == 1)
                mag = 1;
              else if ((mtk > 851.16 && mtk <= 948.16) && old == 1)
                mag = 1;
              else if (mtk > 948.16) {
                mag = 0;
                mag_time = -1.000;
              }
            } else {
              mag = 3;
              mag_time = -1.000;
            }
          } else if (p->dynamo == 1) {
            if (sil) {
              if (mtk <= 851.16) {
                mag = 2;
                if (old <= 1 || old == 3) mag_time = now;
              } else if (mtk > 851.16 && mtk <= 948.16) {
                mag = 1;
                if (old == 0 || old == 3) mag_time = now;
              } else {
                mag = 0;
                mag_time = -1.000;
              }
            } else {
              mag = 3;
              mag_time = -1.000;
            }
          }
          mk.mg[mm1] = (int8_t)mag;
          mk.mgt[mm1] = mag_time;
        }
        // radiogenic heating per kg is the same for every marker of a step
        // (markmart.c:758-785): formed once on the host, times the density
        double mht = 0;
        if (p->is_silicate[mm2])
          mht = hk.sil * mro;
        else if (p->is_iron[mm2])
          mht = hk.fe * mro;
        // scatter weights (markmart.c:840-911)
        W4 w;
        nodewt2(g.gx, g.gy, e.m10, e.m20, x, y, 0, 0, w);
        key = e.m10 * g.Y + e.m20;
        {
          double ddx = 1.0 - dabs(w.wx0 - 0.5), ddy = 1.0 - dabs(w.wy0 - 0.5);
          if (ddx > 0.5 && ddy > 0.5) {
            double swt2 = 1.0 * 4.0 * (ddx - 0.5) * (ddy - 0.5);
            c_nd = mnu * swt2;
            c_wnd = swt2;
          }
        }
        wx0 = w.wx0;
        wx1 = w.wx1;
        wy0 = w.wy0;
        wy1 = w.wy1;
        s_ro = mro;
        s_bb = mbb;
        s_cp = mcp;
        s_kt = mkt;
        s_ht = mht;
        s_tk = mtk;
        s_nu = mnu;
        if (IMM) {  // markmart.c:905: immobile markers carry no temperature
          s_mob = p->immobile[mm2] ? 0.0 : 1.0;
          s_tk *= s_mob;
        }
      }
    }
    // one atomic per run of lanes in the same cell (lanes are cell-sorted in
    // the common case; any order stays correct, just with more runs).  One
    // corner at a time: its ten products are formed, summed over the run and
    // added, so only one corner's values are live (the kernel's registers
    // bound its occupancy).
    const SegScan sc(key, lane);
    const bool put = sc.tail && key >= 0;
    const int m10 = put ? key / g.Y : 0, m20 = put ? key % g.Y : 0;
    const bool in_tile = TILED && (unsigned)(m10 - ti0) < (unsigned)kTile && (unsigned)(m20 - tj0) < (unsigned)kTile;
    auto add = [&](int f, double *acc, int a, int b, double v) {
      if (in_tile)
        atomicAdd(&tacc[f * TA + (m10 - ti0 + a) * TN + (m20 - tj0 + b)], v);
      else
        aadd(&acc[(m10 + a) * g.Y + m20 + b], v);
    };
    {
      const double nd = sc.sum(c_nd), wnd = sc.sum(c_wnd);
      if (put && wnd != 0) {
        add(9, A.nd0, 1, 1, nd);
        add(10, A.wnd, 1, 1, wnd);
      }
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const double ddx = a ? wx1 : wx0, ddy = b ? wy1 : wy0;
        const double swt = 1.0 * ddx * ddy;
        double v8 = 0, v9 = 0;
        if (ddx > 0.5 && ddy > 0.5) {
          const double swt2 = 1.0 * 4.0 * (ddx - 0.5) * (ddy - 0.5);
          v8 = s_nu * swt2;
          v9 = swt2;
        }
        const double r0 = sc.sum(s_ro * swt), r1 = sc.sum(s_bb * swt), r2 = sc.sum(s_cp * swt);
        const double r3 = sc.sum(s_kt * swt), r4 = sc.sum(s_ht * swt), r5 = sc.sum(swt);
        const double r6 = sc.sum(s_tk * swt), r8 = sc.sum(v8), r9 = sc.sum(v9);
        const double r10 = IMM ? sc.sum(s_mob * swt) : 0.0;
        if (put) {
          if (r9 != 0) {
            add(0, A.nu0, a, b, r8);
            add(1, A.wnu, a, b, r9);
          }
          add(2, A.ro0, a, b, r0);
          add(3, A.et0, a, b, r1);
          add(4, A.cp0, a, b, r2);
          add(5, A.kt0, a, b, r3);
          add(6, A.ht0, a, b, r4);
          add(7, A.w0, a, b, r5);
          add(8, A.tk0, a, b, r6);
          if (IMM) add(11, A.w1, a, b, r10);
        }
      }
  }
  if (TILED) {
    __syncthreads();
    double *const dst[12] = {A.nu0, A.wnu, A.ro0, A.et0, A.cp0, A.kt0, A.ht0, A.w0, A.tk0, A.nd0, A.wnd, A.w1};
    for (int k = threadIdx.x; k < TA; k += blockDim.x) {
      const int i = ti0 + k / TN, j = tj0 + k % TN;
      if (i >= g.X || j >= g.Y) continue;
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const double v = tacc[f * TA + k];
        if (v != 0) aadd(&dst[f][i * g.Y + j], v);
      }
    }
  }
}
#define I2E_RONU_INST(TL, IM)                                                                               \
  template __global__ void k_ronu_markers<TL, IM>(const i2e_params *, Grid, NodeF, Markers, int64_t, Acc, HeatK, \
                                                  double, double, int, const double *, const int *, int);
I2E_RONU_INST(false, false)
I2E_RONU_INST(true, false)
I2E_RONU_INST(false, true)
I2E_RONU_INST(true, true)
#undef I2E_RONU_INST

// positive doubles order like their bit patterns
__device__ __forceinline__ void amin_pos(double *a, double v) {
  atomicMin((unsigned long long *)a, (unsigned long long)__double_as_longlong(v));
}
__device__ __forceinline__ void amax_pos(double *a, double v) {
  atomicMax((unsigned long long *)a, (unsigned long long)__double_as_longlong(v));
}

// node normalisation, markmart.c:917-994; mm[0]=numin, mm[1]=numax
// (initialised to +huge / 0 by the caller).  Grid-stride over the nodes with
// a fixed, small grid: the viscosity range is reduced in the block and
// combined with one atomic pair per block (one pair per wavefront on the
// same two addresses serialised to ~95 us at 501^2).
__global__ __launch_bounds__(256) void k_ronu_nodes(const i2e_params *__restrict__ p, NodeF F, Acc A, int N,
                                                   double *mm, int imm) {
  double lmin = 1e300, lmax = 0;
  for (int m3 = blockIdx.x * blockDim.x + threadIdx.x; m3 < N; m3 += gridDim.x * blockDim.x) {
    if (A.wnu[m3]) {
      double v = A.nu0[m3] / A.wnu[m3];
      if (v < p->nubeg) v = p->nubeg;
      if (v > p->nuend) v = p->nuend;
      lmin = fmin(lmin, v);
      lmax = fmax(lmax, v);
      F.nu[m3] = v;
    }
    if (A.wnd[m3]) {
      double v = A.nd0[m3] / A.wnd[m3];
      if (v < p->nubeg) v = p->nubeg;
      if (v > p->nuend) v = p->nuend;
      lmin = fmin(lmin, v);
      lmax = fmax(lmax, v);
      F.nd[m3] = v;
    }
    double w0 = A.w0[m3];
    if (w0) {
      F.ro[m3] = A.ro0[m3] / w0;
      F.et[m3] = A.et0[m3] / w0;
      F.cp[m3] = A.cp0[m3] / w0;
      F.kt[m3] = A.kt0[m3] / w0;
      F.ht[m3] = A.ht0[m3] / w0;
      // T's weight sum is w0 unless some rock type is immobile (markmart.c:905-911)
      if (!imm)
        F.tk[m3] = A.tk0[m3] / w0;
      else if (A.w1[m3])
        F.tk[m3] = A.tk0[m3] / A.w1[m3];
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    lmin = fmin(lmin, __shfl_xor(lmin, o));
    lmax = fmax(lmax, __shfl_xor(lmax, o));
  }
  __shared__ double smin[4], smax[4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smin[w] = lmin;
    smax[w] = lmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
      lmin = fmin(lmin, smin[k]);
      lmax = fmax(lmax, smax[k]);
    }
    if (lmin < 1e300) amin_pos(&mm[0], lmin);
    if (lmax > 0) amax_pos(&mm[1], lmax);
  }
}

// viscosity contrast limit (markmart.c:1002-1009)
__global__ void k_contrast(const i2e_params *__restrict__ p, NodeF F, int N, const double *mm) {
  int m3 = blockIdx.x * blockDim.x + threadIdx.x;
  if (m3 >= N) return;
  double numin = mm[0], numax = mm[1];
  if (numin >= 1e300) numin = 0;  // no weighted node at all: the reference's numin stays 0
  if (p->nucontr >= 1.0)
    numax = numin * p->nucontr;
  else
    numin = numax * p->nucontr;
  double a = F.nu[m3], b = F.nd[m3];
  if (a < numin) a = numin;
  if (a > numax) a = numax;
  if (b < numin) b = numin;
  if (b > numax) b = numax;
  F.nu[m3] = a;
  F.nd[m3] = b;
}

// tkrecalc (heatmart.c:758-826): T = const + sum k T_nbr on BC rows, free
// nodes keep their value.  A row whose neighbours are BC rows is exact once
// they are: the caller runs one pass per level of such chains (a neighbour
// read while another lane writes it gives its old or its new value).
__global__ void k_tbc(NodeF F, BcTab bc, int N) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  int b = bc.bondm[3 * N + m];
  if (!b) return;
  double v = bc.bondv[4 * b];
  for (int n1 = 0; n1 < 3; n1++) {
    int nb = bc.bondn[3 * b + n1];
    if (nb) v += bc.bondv[4 * b + n1 + 1] * F.tk[nb - 1 - 3 * N];
  }
  F.tk[m] = v;
}

// ---------------------------------------------------------------- marker T
// heatmart.c:529-627: subgrid diffusion + scatter of the diffusion change
// TILED: one block per marker tile of the last cell sort, the two node sums
// staged in LDS as in k_ronu_markers
template <bool TILED>
__global__ __launch_bounds__(256) void k_heat_markers_a(const i2e_params *__restrict__ p, Grid g, NodeF F,
                                                       const double *__restrict__ tk2, Markers mk, int64_t M,
                                                       double timestep, double *sol0, double *sol1,
                                                       const int *__restrict__ toff, int ntj) {
  constexpr int TN = kTile + 1, TA = TN * TN;
  __shared__ double tacc[TILED ? 2 * TA : 1];
  const int lane = threadIdx.x & 63;
  int ti0 = 0, tj0 = 0;
  int64_t lo, hi, stride;
  if (TILED) {
    const int kk = kTile * kTile, T = blockIdx.x;
    lo = T ? toff[T * kk - 1] : 0;
    hi = toff[T * kk + kk - 1];
    stride = blockDim.x;
    ti0 = (T / ntj) * kTile;
    tj0 = (T % ntj) * kTile;
    for (int k = threadIdx.x; k < 2 * TA; k += blockDim.x) tacc[k] = 0;
    __syncthreads();
  } else {
    lo = (int64_t)blockIdx.x * blockDim.x;
    hi = M;
    stride = (int64_t)gridDim.x * blockDim.x;
  }
  // whole wavefronts iterate together: the scatter reduces across lanes
  for (int64_t base = lo; base < hi; base += stride) {
    const int64_t m3 = base + threadIdx.x;
    double s0[4] = {0, 0, 0, 0}, s1[4] = {0, 0, 0, 0};
    int key = -1;
    if (m3 < hi) {
      float fx = mk.x[m3], fy = mk.y[m3];
      if (fx > 0 && fy > 0 && (double)fx < g.xsize && (double)fy < g.ysize && mk.t[m3] < 50) {
        double x = fx, y = fy;
        int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
        W4 w;
        nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
        double e2 = isum(w, g.Y, F.tk), e3 = isum(w, g.Y, tk2);
        float fk = mk.k[m3];
        if (fk <= 0) fk = (float)e3;
        if (p->heatdif) {
          double dtk = e3 - (double)fk;
          double celdx = g.gx[m10 + 1] - g.gx[m10], celdy = g.gy[m20 + 1] - g.gy[m20];
          double mkt = isum(w, g.Y, F.kt), mcp = isum(w, g.Y, F.cp), mro = isum(w, g.Y, F.ro);
          double dtk1 = -p->heatdif * mkt / mro / mcp * 2.0 * (1.0 / celdx / celdx + 1.0 / celdy / celdy) * timestep;
          if (dtk1 < -150.0) dtk1 = -150.0;
          dtk1 = dtk * (1.0 - exp(dtk1));
          fk += (float)dtk1;
          key = m10 * g.Y + m20;
#pragma unroll
          for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
              double ival = (a ? w.wx1 : w.wx0) * (b ? w.wy1 : w.wy0) * 1.0;
              s0[a * 2 + b] = ival;
              s1[a * 2 + b] = dtk1 * ival;
            }
        }
        fk += (float)(e2 - e3);
        mk.k[m3] = fk;
      }
    }
    if (!p->heatdif) continue;
    // one atomic per run of lanes in the same cell
    const SegScan sc(key, lane);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      s0[c] = sc.sum(s0[c]);
      s1[c] = sc.sum(s1[c]);
    }
    if (sc.tail && key >= 0) {
      const int m10 = key / g.Y, m20 = key % g.Y;
      const bool in_tile =
          TILED && (unsigned)(m10 - ti0) < (unsigned)kTile && (unsigned)(m20 - tj0) < (unsigned)kTile;
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          if (in_tile) {
            const int l = (m10 - ti0 + a) * TN + (m20 - tj0 + b);
            atomicAdd(&tacc[l], s0[a * 2 + b]);
            atomicAdd(&tacc[TA + l], s1[a * 2 + b]);
          } else {
            const int m4 = (m10 + a) * g.Y + m20 + b;
            aadd(&sol0[m4], s0[a * 2 + b]);
            aadd(&sol1[m4], s1[a * 2 + b]);
          }
        }
    }
  }
  if (TILED && p->heatdif) {
    __syncthreads();
    for (int k = threadIdx.x; k < TA; k += blockDim.x) {
      const int i = ti0 + k / TN, j = tj0 + k % TN;
      if (i >= g.X || j >= g.Y) continue;
      const double a = tacc[k], b = tacc[TA + k];
      if (a != 0) aadd(&sol0[i * g.Y + j], a);
      if (b != 0) aadd(&sol1[i * g.Y + j], b);
    }
  }
}
template __global__ void k_heat_markers_a<false>(const i2e_params *, Grid, NodeF, const double *, Markers, int64_t,
                                                 double, double *, double *, const int *, int);
template __global__ void k_heat_markers_a<true>(const i2e_params *, Grid, NodeF, const double *, Markers, int64_t,
                                                double, double *, double *, const int *, int);

__global__ void k_heat_nodes_norm(int N, const double *sol0, double *sol1) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < N && sol0[m]) sol1[m] /= sol0[m];
}

// heatmart.c:635-688: antidiffusion
__global__ void k_heat_markers_b(Grid g, Markers mk, int64_t M, const double *__restrict__ sol1) {
  for (int64_t m3 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m3 < M; m3 += (int64_t)gridDim.x * blockDim.x) {
    float fx = mk.x[m3], fy = mk.y[m3];
    if (!(fx > 0 && fy > 0 && (double)fx < g.xsize && (double)fy < g.ysize && mk.t[m3] < 50)) continue;
    double x = fx, y = fy;
    int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
    W4 w;
    nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
    float fk = mk.k[m3];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        int m4 = (m10 + a) * g.Y + m20 + b;
        double ival = (a ? w.wx1 : w.wx0) * (b ? w.wy1 : w.wy0);
        double s = sol1[m4];
        if (s) fk -= (float)(s * ival);
      }
    mk.k[m3] = fk;
  }
}

// per-cell marker-type histogram (cell = m1serch,m2serch of the position)
// ---------------------------------------------------------------- body statistics
// The node pass (impact.c:3064-3076): the largest node temperature (at least
// the ambient) and the sum and count of the body's nodes (tk != ambient).
__global__ __launch_bounds__(256) void k_hyd_nodes(int N, const double *__restrict__ tk, double tamb,
                                                  unsigned long long *acc, double *dacc) {
  double mx = tamb, sm = 0;
  unsigned cnt = 0;
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < N; m += gridDim.x * blockDim.x) {
    const double t = tk[m];
    mx = fmax(mx, t);
    if (t != tamb) {
      sm += t;
      cnt++;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, o));
    sm += __shfl_xor(sm, o);
    cnt += __shfl_xor(cnt, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (cnt) {
      atomicAdd(&acc[17], (unsigned long long)cnt);
      aadd(&dacc[1], sm);
    }
    // node temperatures are positive: their bit patterns order like them
    atomicMax((unsigned long long *)&dacc[2], (unsigned long long)__double_as_longlong(mx));
  }
}

// The marker pass (impact.c:3078-3206): the rheology switch between rock+ice
// (5) and hydrous silicate (6) by the peak temperature, then the counts of
// the silicate markers (5, 6, 25, 26) by type and peak temperature.
__global__ __launch_bounds__(256) void k_hyd_markers(Markers mk, int64_t M, double t_sol, double t_liq, int hydrate,
                                                    const double *__restrict__ maxtk, double tamb,
                                                    unsigned long long *acc, double *dacc) {
  enum { SOL, LIQ, HYDROUS, PRIMITIVE, N2CO2, COCL, H2O, PH1, PH2, PH3, PH4, PERCO, MELT1, MELT2, TOT, TOOHOT, NC };
  unsigned c[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) c[k] = 0;
  unsigned tmaxb = 0;  // bits of the largest silicate marker T above the ambient (positive floats)
  double tsum = 0;
  const double mxt = *maxtk;
  if (hydrate) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
      int t = mk.t[i];
      const double tm = (double)mk.tmax[i];
      if (t == 5 && tm >= 273.15 && tm < 1223.15) t = 6;
      if (t == 6 && tm >= 1223.15) t = 5;
      if (t != mk.t[i]) mk.t[i] = (int8_t)t;
      if (t == 5 || t == 6) c[SOL]++;
      if (t == 25 || t == 26) c[LIQ]++;
      if (t == 5 || t == 6 || t == 25 || t == 26) {
        c[PRIMITIVE] += tm < 273.15;
        c[HYDROUS] += tm >= 273.15 && tm < 1223.15;
        c[N2CO2] += tm < 1073.15;
        c[COCL] += tm < (0.10 * (t_liq - t_sol) + t_sol);
        c[H2O] += tm < 1223.15;
        c[PH1] += tm >= 273.15 && tm < 773.15;
        c[PH2] += tm >= 773.15 && tm < 873.15;
        c[PH3] += tm >= 873.15 && tm < 973.15;
        c[PH4] += tm >= 973.15 && tm < 1223.15;
        c[PERCO] += tm >= 1273.15;
        c[MELT1] += tm >= t_sol;
        c[MELT2] += tm >= (0.40 * (t_liq - t_sol) + t_sol);
        c[TOT]++;
        const float k = mk.k[i];
        if ((double)k > tamb && __float_as_uint(k) > tmaxb) tmaxb = __float_as_uint(k);
        tsum += (double)k;
        if ((double)k / mxt > 1.1) c[TOOHOT]++;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NC; k++)
    for (int o = 32; o > 0; o >>= 1) c[k] += __shfl_xor(c[k], o);
  for (int o = 32; o > 0; o >>= 1) {
    tsum += __shfl_xor(tsum, o);
    tmaxb = max(tmaxb, (unsigned)__shfl_xor((int)tmaxb, o));
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < NC; k++)
      if (c[k]) atomicAdd(&acc[k], (unsigned long long)c[k]);
    if (tmaxb) atomicMax(&acc[16], (unsigned long long)tmaxb);
    if (tsum != 0) aadd(&dacc[0], tsum);
  }
}

__global__ void k_type_counts(Grid g, Markers mk, int64_t M, int ntypes, int *cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    float fx = mk.x[i], fy = mk.y[i];
    int t = mk.t[i];
    if (!(fx > 0 && fy > 0 && (double)fx < g.xsize && (double)fy < g.ysize)) continue;
    if (t < 0 || t >= ntypes) continue;
    int c = serch(g.gx, g.X, (double)fx) * (g.Y - 1) + serch(g.gy, g.Y, (double)fy);
    atomicAdd(&cnt[c * ntypes + t], 1);
  }
}

}  // namespace i2e
