// postproc.hip -- post-processing on the device: the visualisation grid of
// i2jmart's saver (loadjmart.c:619-1197) and the text files saver() and
// markersave() write; the olivine grain-growth update of grain() (grain.c).
//
// saver() walks the markers in index order and adds, in float32, each
// marker's value times its bilinear weight and the weight itself to the four
// pixels around it; a pixel's sum in the reference is therefore the float
// sum of its contributions in marker order.  Here the markers are binned by
// the pixel cell they fall in (a counting sort, each cell's list then put in
// marker order), and one thread per pixel merges the lists of its four cells
// by marker index and accumulates exactly the reference's float additions
// in that order: the grid equals the serial one bit for bit wherever the
// values are formed the same way (the device's log and pow may round a
// value the host's would not).
#include <stdio.h>
#include <string.h>

#include <functional>
#include <vector>

#include "ctx.hpp"
#include "kernels.hpp"

namespace i2e {

struct VisGeo {
  double xminx, xmaxx, yminy, ymaxy, xminx0, yminy0, xresx, yresy;
  int xresol, yresol;
  int64_t markstp;
};

// allinter1 (loadjmart.c:1327-1520, intermod 0): the interpolated values saver() reads
struct E1 {
  double t, nu, ro, kt, ht, exy, sxy, exysxy, esp, exx, sxx, eyy, syy, pr, exxsxx, eyysyy;
};
__device__ void allinter1_d(const Grid &g, const NodeF &F, const double *esp, double x, double y, E1 &e) {
  W4 w;
  if (x < 0) x = 0;
  else if (x > g.xsize) x = g.xsize;
  if (y < 0) y = 0;
  else if (y > g.ysize) y = g.ysize;
  const int m10 = serch(g.gx, g.X, x), m20 = serch(g.gy, g.Y, y);
  nodewt2(g.gx, g.gy, m10, m20, x, y, 0, 0, w);
  e.t = isum(w, g.Y, F.tk);
  e.nu = isum(w, g.Y, F.nu);
  e.ro = isum(w, g.Y, F.ro);
  e.kt = isum(w, g.Y, F.kt);
  e.ht = isum(w, g.Y, F.ht);
  nodewt2(g.gx, g.gy, bn_i0(m10, g.X), bn_i0(m20, g.Y), x, y, 0, 0, w);
  e.exy = isum(w, g.Y, F.exy);
  e.sxy = isum(w, g.Y, F.sxy);
  e.exysxy = isum2(w, g.Y, F.exy, F.sxy);
  e.esp = isum(w, g.Y, esp);
  nodewt2(g.gx, g.gy, ce_i0(g.gx, m10, x, g.X), ce_i0(g.gy, m20, y, g.Y), x, y, -1, -1, w);
  e.exx = isum(w, g.Y, F.exx);
  e.sxx = isum(w, g.Y, F.sxx);
  e.eyy = isum(w, g.Y, F.eyy);
  e.syy = isum(w, g.Y, F.syy);
  e.pr = isum(w, g.Y, F.pr);
  e.exxsxx = isum2(w, g.Y, F.exx, F.sxx);
  e.eyysyy = isum2(w, g.Y, F.eyy, F.syy);
}

// spncalc (loadjmart.c:1523-1617, stoksfd 0): rotation rate at the interior
// basic nodes, 0 elsewhere
__global__ void k_vis_esp(Grid g, NodeF F, double *esp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  const int X = g.X, Y = g.Y;
  if (j >= Y) return;
  double r = 0;
  if (i >= 1 && i <= X - 2 && j >= 1 && j <= Y - 2) {
    const int j0 = j - 1, j1 = j > Y - 2 ? Y - 2 : j, i0 = i - 1, i1 = i > X - 2 ? X - 2 : i;
    double d0, d1, a = 0, b = 0;
    fd2_deriv((g.gy[j0] + g.gy[j0 + 1]) / 2.0, (g.gy[j1] + g.gy[j1 + 1]) / 2.0, d0, d1);
    a += F.vx[i * Y + j0] * d0 / 2.0;
    a += F.vx[i * Y + j1] * d1 / 2.0;
    fd2_deriv((g.gx[i0] + g.gx[i0 + 1]) / 2.0, (g.gx[i1] + g.gx[i1 + 1]) / 2.0, d0, d1);
    b += F.vy[i0 * Y + j] * d0 / 2.0;
    b += F.vy[i1 * Y + j] * d1 / 2.0;
    r = b - a;
  }
  esp[i * Y + j] = r;
}

// the value saver() gives a marker (loadjmart.c:893-998; markht is not
// carried, so heat generation is the shear heating)
__device__ double marker_value(const i2e_params *p, int type, int t, double mtk, float markv, float por,
                               const E1 &e) {
  const double mpb = e.pr * 1e-5;
  double mnu = markv;
  if (mnu < p->nubeg) mnu = p->nubeg;
  if (mnu > p->nuend) mnu = p->nuend;
  const i2e_rock &r = p->rock[t];
  double mro = r.ro;
  if (p->densimod != 0)
    mro = r.ro * (1.0 - r.bb * (mtk - 298.15)) * (1.0 + r.aa * (mpb - 1.0) * 1e-3) * (1.000 - (double)por);
  double mkt = (r.kt + r.kf / (mtk + 77.0)) * exp(r.kp * mpb);
  if (r.kt < 0) mkt = -r.kt / (1.0 + r.kf * (mtk - r.kp) / r.kp);
  const double mht = 0.0;
  double ival = 0;
  switch (type) {
    case 1: ival = mtk; break;
    case 2: ival = log(mnu) / log(10.0); break;
    case 4:
      ival = 2.0 * e.exysxy + e.exxsxx + e.eyysyy;
      if (ival < 1e-30) ival = 1e-30;
      ival = log(ival) / log(10.0);
      break;
    case 6: ival = mro; break;
    case 7: ival = mkt; break;
    case 8: ival = mht + 2.0 * e.exysxy + e.exxsxx + e.eyysyy; break;
    case 12:
      ival = pow(e.exy * e.exy + 0.5 * (e.exx * e.exx + e.eyy * e.eyy), 0.5);
      if (ival < 1e-30) ival = 1e-30;
      ival = e.esp / ival;
      break;
  }
  return ival;
}

// per marker slot: the pixel cell key m1 (yresol + 1) + m2 of the sampled
// window markers (-1 otherwise), the fractional offsets, the value (or the
// rock type), the melt fraction by marker index; cell counts
__global__ void k_vis_markers(const i2e_params *__restrict__ p, Grid g, NodeF F, const double *esp, Markers mk,
                              int64_t M, const int *orig, VisGeo v, int type, int *key, int *mix, double *ea_o,
                              double *na_o, double *val_o, int *cnt, int *cmm, double *melt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const int o = orig ? orig[i] : (int)i;
    key[i] = -1;
    if (o % v.markstp) continue;
    const float fx = mk.x[i], fy = mk.y[i];
    const int t = mk.t[i];
    if (!((double)fx > v.xminx && (double)fx < v.xmaxx && (double)fy > v.yminy && (double)fy < v.ymaxy && t < 50))
      continue;
    if (type == 5) {
      atomicMin(&cmm[0], t);
      atomicMax(&cmm[1], t);
      val_o[i] = (double)t;
    } else {
      E1 e;
      allinter1_d(g, F, esp, (double)fx, (double)fy, e);
      float fk = mk.k[i];
      if (fk < 1.0) fk = (float)e.t;
      const double mtk = (double)fk;
      if (melt) {
        double xm = 0;
        const int base = t < 20 ? t : t - 20;
        if (p->melt_on && base >= 0 && base < I2E_MAXTYPES && p->meltable[base]) {
          xm = (mtk - p->melt_ts) / (p->melt_tl - p->melt_ts);
          if (xm < 0.0) xm = 0.0;
          if (xm > 1.0) xm = 1.0;
        }
        melt[o] = xm;
      }
      val_o[i] = marker_value(p, type, t, mtk, mk.v[i], mk.por[i], e);
    }
    double ea = ((double)(fx)-v.xminx0) / v.xresx + 1.0;
    int m1 = (int)(ea);
    if (m1 < 0) m1 = 0;
    if (m1 > v.xresol) m1 = v.xresol;
    ea -= (double)(m1);
    double na = ((double)(fy)-v.yminy0) / v.yresy + 1.0;
    int m2 = (int)(na);
    if (m2 < 0) m2 = 0;
    if (m2 > v.yresol) m2 = v.yresol;
    na -= (double)(m2);
    const int k = m1 * (v.yresol + 1) + m2;
    key[i] = k;
    mix[i] = o;
    ea_o[i] = ea;
    na_o[i] = na;
    atomicAdd(&cnt[k], 1);
  }
}

// the sampled markers into their cells (cur: the exclusive scan of the
// counts, advanced atomically)
__global__ void k_vis_place(int64_t M, const int *key, int *cur, int *list) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const int k = key[i];
    if (k >= 0) list[atomicAdd(&cur[k], 1)] = (int)i;
  }
}

// each cell's list in marker-index order (lists hold a few to a few dozen)
__global__ void k_vis_cellsort(int C, const int *off, const int *cnt, const int *mix, int *list) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  int *l = list + off[c];
  const int n = cnt[c];
  for (int a = 1; a < n; a++) {
    const int s = l[a], ks = mix[s];
    int b = a - 1;
    while (b >= 0 && mix[l[b]] > ks) {
      l[b + 1] = l[b];
      b--;
    }
    l[b + 1] = s;
  }
}

// the four cells that feed pixel (p1, p2) and the corner of the marker's
// cell the pixel is: cell (p1+1, p2+1) corner 0, (p1+1, p2) corner 1,
// (p1, p2+1) corner 2, (p1, p2) corner 3 (loadjmart.c:1015-1020)
struct Merge4 {
  const int *l[4];
  int n[4], q[4];
};
__device__ __forceinline__ void merge_init(Merge4 &mg, const VisGeo &v, const int *off, const int *cnt,
                                           const int *list, int p1, int p2) {
  const int Yc = v.yresol + 1;
  const int cells[4] = {(p1 + 1) * Yc + p2 + 1, (p1 + 1) * Yc + p2, p1 * Yc + p2 + 1, p1 * Yc + p2};
  for (int k = 0; k < 4; k++) {
    mg.l[k] = list + off[cells[k]];
    mg.n[k] = cnt[cells[k]];
    mg.q[k] = 0;
  }
}
// next marker slot in index order and its corner; false when done
__device__ __forceinline__ bool merge_next(Merge4 &mg, const int *mix, int &slot, int &corner) {
  int best = -1, bo = 0x7fffffff;
  for (int k = 0; k < 4; k++)
    if (mg.q[k] < mg.n[k]) {
      const int o = mix[mg.l[k][mg.q[k]]];
      if (o < bo) {
        bo = o;
        best = k;
      }
    }
  if (best < 0) return false;
  slot = mg.l[best][mg.q[best]++];
  corner = best;
  return true;
}
__device__ __forceinline__ double corner_wt(int corner, double ea, double na) {
  switch (corner) {
    case 0: return (1.0 - ea) * (1.0 - na);
    case 1: return (1.0 - ea) * na;
    case 2: return ea * (1.0 - na);
    default: return ea * na;
  }
}
__device__ __forceinline__ double corner_val(int corner, double ival, double ea, double na) {
  switch (corner) {
    case 0: return ival * (1.0 - ea) * (1.0 - na);
    case 1: return ival * (1.0 - ea) * na;
    case 2: return ival * ea * (1.0 - na);
    default: return ival * ea * na;
  }
}

// transport-property pixels (loadjmart.c:1010-1121)
__global__ void k_vis_pixels(const i2e_params *__restrict__ p, Grid g, NodeF F, const double *esp, VisGeo v,
                             int type, const int *off, const int *cnt, const int *list, const int *mix,
                             const double *ea, const double *na, const double *val, float *out) {
  const int p2 = blockIdx.x * blockDim.x + threadIdx.x, p1 = blockIdx.y;
  if (p2 >= v.yresol) return;
  Merge4 mg;
  merge_init(mg, v, off, cnt, list, p1, p2);
  float s = 0, w = 0;
  int slot, corner;
  while (merge_next(mg, mix, slot, corner)) {
    s += (float)(corner_val(corner, val[slot], ea[slot], na[slot]));
    w += (float)(corner_wt(corner, ea[slot], na[slot]));
  }
  float r;
  if (w) {
    r = s / w;
  } else {
    E1 e;
    allinter1_d(g, F, esp, v.xminx0 + v.xresx * ((double)(p1)), v.yminy0 + v.yresy * ((double)(p2)), e);
    r = 0;
    switch (type) {
      case 1: r = (float)(e.t); break;
      case 2: r = (float)(log(e.nu) / log(10.0)); break;
      case 4:
        r = (float)(2.0 * e.exysxy + e.exxsxx + e.eyysyy);
        if (r < 1e-30) r = 1e-30;
        r = (float)(log(r) / log(10.0));
        break;
      case 6: r = (float)(e.ro); break;
      case 7: r = (float)(e.kt); break;
      case 8: r = (float)(e.ht + 2.0 * e.exysxy + e.exxsxx + e.eyysyy); break;
      case 12: {
        double iv = pow(e.exy * e.exy + 0.5 * (e.exx * e.exx + e.eyy * e.eyy), 0.5);
        if (iv < 1e-30) iv = 1e-30;
        r = e.esp / iv;
        break;
      }
    }
  }
  out[p1 * v.yresol + p2] = r;
}

// rock-type pixels (loadjmart.c:722-780): per type in order, the float sum
// of the weights of that type's markers (continuing a loser's sum, as the
// reference leaves it uncleared) against the best so far; -1 if none
__global__ void k_vis_rock(VisGeo v, const int *cmm, const int *off, const int *cnt, const int *list, const int *mix,
                           const double *ea, const double *na, const double *val, float *best, int *lin) {
  const int p2 = blockIdx.x * blockDim.x + threadIdx.x, p1 = blockIdx.y;
  if (p2 >= v.yresol) return;
  float b = 0, acc = 0;
  int l = 0;
  for (int ccur = cmm[0]; ccur <= cmm[1]; ccur++) {
    Merge4 mg;
    merge_init(mg, v, off, cnt, list, p1, p2);
    int slot, corner;
    while (merge_next(mg, mix, slot, corner))
      if ((int)val[slot] == ccur) acc += (float)(corner_wt(corner, ea[slot], na[slot]));
    if (acc > b) {
      b = acc;
      l = ccur;
      acc = 0;
    }
  }
  const int m = p1 * v.yresol + p2;
  best[m] = b;
  lin[m] = b ? l : -1;
}

// empty pixels from their neighbours within gridmod (loadjmart.c:792-868):
// the nearest ring holding typed pixels bounds the types, each type's
// distance-weighted sum of the neighbours' winning weights (again continuing
// a loser's sum) against the best
__global__ void k_vis_fill(VisGeo v, int gridmod, const float *best, const int *lin, int *lin2) {
  const int m2 = blockIdx.x * blockDim.x + threadIdx.x, m1 = blockIdx.y;
  const int X = v.xresol, Y = v.yresol;
  if (m2 >= Y) return;
  const int m3 = m1 * Y + m2;
  lin2[m3] = lin[m3];
  if (lin[m3] != -1) return;
  int r = 1, yn = 0, cmin = 1000, cmax = 0;
  do {
    cmin = 1000, cmax = 0;
    for (int m5 = m1 - r; m5 <= m1 + r; m5++)
      for (int m6 = m2 - r; m6 <= m2 + r; m6++)
        if (m5 >= 0 && m5 < X && m6 >= 0 && m6 < Y) {
          const int c = lin[m5 * Y + m6];
          if (c != -1) {
            yn = 1;
            if (cmin > c) cmin = c;
            if (cmax < c) cmax = c;
          }
        }
    r++;
  } while (!yn && r <= gridmod);
  if (!yn) return;
  float own = 0, acc = 0;
  int got = 0;
  for (int ccur = cmin; ccur <= cmax; ccur++) {
    for (int m5 = m1 - r + 1; m5 < m1 + r; m5++)
      for (int m6 = m2 - r + 1; m6 < m2 + r; m6++)
        if (m5 >= 0 && m5 < X && m6 >= 0 && m6 < Y) {
          const int m7 = m5 * Y + m6;
          if (lin[m7] == ccur) {
            const double ea = dabs(((double)(m5 - m1)) / (double)(r)), na = dabs(((double)(m6 - m2)) / (double)(r));
            acc += (float)((1.0 - ea) * (1.0 - na)) * best[m7];
          }
        }
    if (acc > own) {
      own = acc;
      got = ccur;
      acc = 0;
    }
  }
  if (own) lin2[m3] = got;
}

// grain() (grain.c:4-156): olivine grain growth d^n <- d^n + k0 exp(-E/(R T))
// dt per marker, after the first impact (or from the start, growth_model
// 2), and the pallasite emplacement in a mantle sector during the impact's
// time window (types 6 -> 5, 26 -> 25, grain size reset).  Markers in slot
// order; gr holds the grain sizes by slot.
struct GrainK {
  int growth_model;
  double gr_init, t_after, t_impact, t_window_end, t_start, timestep, xsize, ysize;
};
__global__ void k_grain(GrainK q, Markers mk, int64_t M, float *gr) {
  const double n_grain = 3.70, E_grain = 1.0098e+5, k0_grain = 3.2033, n_oliv = 2.250, E_oliv = 5.71674e+5,
               k0_oliv = 1.06593e+16, pival = 3.141592654, radius_core = 1.000e+5, radius_planet = 2.000e+5,
               ref_angle = 45.000, dev_angle = 10.000;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    int t = mk.t[i];
    float g = gr[i];
    const float k = mk.k[i];
    if ((q.growth_model == 1 || q.growth_model == 2) && q.t_after >= q.t_impact && q.t_after <= q.t_window_end) {
      const double dist_x = mk.x[i] - (q.xsize / 2.000), dist_y = mk.y[i] - (q.ysize / 2.000);
      const double d = pow((dist_x * dist_x + dist_y * dist_y), 0.500);
      if (d >= radius_core && d <= radius_planet && dist_x > 0.000 && dist_y < 0.000) {
        const double ang = asin(dist_x / d) / pival * 180.000;
        if (ang >= ref_angle - dev_angle && ang <= ref_angle + dev_angle) {
          if (t == 6) t = 5;
          if (t == 26) t = 25;
          g = (float)(q.gr_init);
        }
      }
    }
    auto grow = [&](double n, double k0, double E) {
      const double a = pow(((double)(g)), n), b = k0 * exp(-E / (8.3145 * k)) * q.timestep;
      g = (float)(pow((a + b), (1.000 / n)));
    };
    if (q.t_after >= q.t_impact) {
      if (t == 5 && k >= 1073.000) grow(n_grain, k0_grain, E_grain);
      if (t == 25 && k < 2103.000) grow(n_grain, k0_grain, E_grain);
    }
    if ((q.growth_model == 1 && q.t_after >= q.t_impact) || (q.growth_model == 2 && q.t_after >= q.t_start)) {
      if (t == 6) grow(n_oliv, k0_oliv, E_oliv);
      if (t == 26 && k < 2103.000) grow(n_oliv, k0_oliv, E_oliv);
      if ((t == 25 || t == 26) && k >= 2103.000) g = (float)(q.gr_init);
    }
    mk.t[i] = (int8_t)t;
    gr[i] = g;
  }
}

}  // namespace i2e

using namespace i2e;

static inline dim3 grid2(int X, int Y) { return dim3((Y + 255) / 256, X); }
static inline unsigned mgrid_v(int64_t M) { return nblk(M, 256) < 65535u ? nblk(M, 256) : 65535u; }

extern "C" {

int i2e_visual(i2e_ctx *c, const i2e_visual_spec *sp, float *val, int32_t *rock, double *melt) {
  if (!c || !sp || !c->setup_done) return I2E_ERR_ARG;
  const int type = sp->type;
  if (type < 1 || type > 12 || type == 3 || (type >= 9 && type <= 11)) {
    set_error("visual: output type %d is not one saver() computes (1, 2, 4-8, 12)", type);
    return I2E_ERR_UNSUPPORTED;
  }
  if (c->P.densimod == 2) {
    set_error("visual: densimod 2 (thermodynamic database) is not built");
    return I2E_ERR_UNSUPPORTED;
  }
  const int64_t xr = sp->xresol, yr = sp->yresol, nres = xr * yr;
  if (xr < 2 || yr < 2 || nres * 2 > 40000000 || (type == 5 ? !rock : !val)) {
    set_error("visual: resolution %lld x %lld outside 2..(val0[] holds 2 x 2e7), or no output array",
              (long long)xr, (long long)yr);
    return I2E_ERR_ARG;
  }
  const int X = c->X, Y = c->Y, N = c->N;
  const int64_t M = c->M;
  const Grid g = c->grid();
  // the window of loadjmart.c:670-708 (one chunk)
  VisGeo v;
  double xb = sp->xbeg, xe = sp->xend, yb = sp->ybeg, ye = sp->yend;
  if (xb < 0) xb = 0;
  if (xe > 1.0) xe = 1.0;
  if (yb < 0) yb = 0;
  if (ye > 1.0) ye = 1.0;
  v.xminx0 = xb * c->P.xsize;
  v.xresx = (xe * c->P.xsize - v.xminx0) / (double)(xr - 1);
  v.yminy0 = yb * c->P.ysize;
  v.yresy = (ye * c->P.ysize - v.yminy0) / (double)(yr - 1);
  v.yminy = v.yminy0 - v.yresy;
  if (v.yminy < 0) v.yminy = 0;
  v.ymaxy = ye * c->P.ysize + v.yresy;
  if (v.ymaxy > c->P.ysize) v.ymaxy = c->P.ysize;
  v.xminx = v.xminx0 - v.xresx;
  if (v.xminx < 0) v.xminx = 0;
  v.xmaxx = xb * c->P.xsize + v.xresx * (double)xr;
  if (v.xmaxx > c->P.xsize) v.xmaxx = c->P.xsize;
  v.xresol = (int)xr;
  v.yresol = (int)yr;
  v.markstp = (int64_t)(0.1 * (double)M / (double)nres);
  if (v.markstp < 1) v.markstp = 1;
  const int C = (int)((xr + 1) * (yr + 1)), nb = (C + SB - 1) / SB;
  const size_t MM = M ? (size_t)M : 1;
  // device buffers of this call
  DBuf<double> esp, mea, mna, mval, dmelt;
  DBuf<int> key, mix, cnt, off, cur, list, bsum, cmm, lin, lin2;
  DBuf<float> fout;
  struct Rel {
    std::vector<std::function<void()>> f;
    ~Rel() {
      for (auto &x : f) x();
    }
  } rel;
  rel.f = {[&] { esp.release(); }, [&] { mea.release(); }, [&] { mna.release(); }, [&] { mval.release(); },
           [&] { dmelt.release(); }, [&] { key.release(); }, [&] { mix.release(); }, [&] { cnt.release(); },
           [&] { off.release(); }, [&] { cur.release(); }, [&] { list.release(); }, [&] { bsum.release(); },
           [&] { cmm.release(); }, [&] { lin.release(); }, [&] { lin2.release(); }, [&] { fout.release(); }};
  ST_OK(esp.alloc(N));
  ST_OK(mea.alloc(MM));
  ST_OK(mna.alloc(MM));
  ST_OK(mval.alloc(MM));
  ST_OK(key.alloc(MM));
  ST_OK(mix.alloc(MM));
  ST_OK(list.alloc(MM));
  ST_OK(cnt.alloc(C));
  ST_OK(off.alloc(C));
  ST_OK(cur.alloc(C));
  ST_OK(bsum.alloc(nb));
  ST_OK(cmm.alloc(2));
  ST_OK(fout.alloc(nres));
  ST_OK(lin.alloc(nres));
  ST_OK(lin2.alloc(nres));
  if (melt && type != 5) {
    ST_OK(dmelt.alloc(MM));
    HIP_OK(hipMemsetAsync(dmelt.p, 0, sizeof(double) * MM, c->s));
  }
  const int cinit[2] = {1000, 0};
  HIP_OK(hipMemcpyAsync(cmm.p, cinit, sizeof cinit, hipMemcpyHostToDevice, c->s));
  HIP_OK(hipMemsetAsync(cnt.p, 0, sizeof(int) * C, c->s));
  hipLaunchKernelGGL(k_vis_esp, grid2(X, Y), dim3(256), 0, c->s, g, c->F, esp.p);
  const unsigned gm = mgrid_v(M ? M : 1);
  if (M)
    hipLaunchKernelGGL(k_vis_markers, dim3(gm), dim3(256), 0, c->s, c->dP, g, c->F, (const double *)esp.p, c->mk, M,
                       c->permuted ? (const int *)c->morig.p : nullptr, v, type, key.p, mix.p, mea.p, mna.p, mval.p,
                       cnt.p, cmm.p, dmelt.p);
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, c->s, C, (const int *)cnt.p, off.p, bsum.p);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, c->s, nb, bsum.p);
  hipLaunchKernelGGL(k_scan_add, dim3(nblk(C, 256)), dim3(256), 0, c->s, C, (const int *)bsum.p, off.p);
  HIP_OK(hipMemcpyAsync(cur.p, off.p, sizeof(int) * C, hipMemcpyDeviceToDevice, c->s));
  if (M) hipLaunchKernelGGL(k_vis_place, dim3(gm), dim3(256), 0, c->s, M, (const int *)key.p, cur.p, list.p);
  hipLaunchKernelGGL(k_vis_cellsort, dim3(nblk(C, 256)), dim3(256), 0, c->s, C, (const int *)off.p,
                     (const int *)cnt.p, (const int *)mix.p, list.p);
  if (type != 5) {
    hipLaunchKernelGGL(k_vis_pixels, grid2((int)xr, (int)yr), dim3(256), 0, c->s, c->dP, g, c->F,
                       (const double *)esp.p, v, type, (const int *)off.p, (const int *)cnt.p, (const int *)list.p,
                       (const int *)mix.p, (const double *)mea.p, (const double *)mna.p, (const double *)mval.p,
                       fout.p);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(val, fout.p, sizeof(float) * nres, hipMemcpyDeviceToHost, c->s));
  } else {
    hipLaunchKernelGGL(k_vis_rock, grid2((int)xr, (int)yr), dim3(256), 0, c->s, v, (const int *)cmm.p,
                       (const int *)off.p, (const int *)cnt.p, (const int *)list.p, (const int *)mix.p,
                       (const double *)mea.p, (const double *)mna.p, (const double *)mval.p, fout.p, lin.p);
    int *res = lin.p;
    if (sp->gridmod > 0) {
      hipLaunchKernelGGL(k_vis_fill, grid2((int)xr, (int)yr), dim3(256), 0, c->s, v, sp->gridmod,
                         (const float *)fout.p, (const int *)lin.p, lin2.p);
      res = lin2.p;
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(rock, res, sizeof(int32_t) * nres, hipMemcpyDeviceToHost, c->s));
  }
  if (melt) {
    if (type != 5 && M)
      HIP_OK(hipMemcpyAsync(melt, dmelt.p, sizeof(double) * M, hipMemcpyDeviceToHost, c->s));
    else if (M)
      memset(melt, 0, sizeof(double) * M);
  }
  HIP_OK(hipStreamSynchronize(c->s));
  return I2E_OK;
}

int i2e_grain(i2e_ctx *c, const i2e_grain_params *gp, float *markgr) {
  if (!c || !gp || (c->M && !markgr)) return I2E_ERR_ARG;
  const int64_t M = c->M;
  if (!M) return I2E_OK;
  GrainK q;
  q.growth_model = gp->growth_model;
  q.gr_init = gp->gr_init;
  q.t_after = c->timesum + c->timestep;
  // the reference's products, left to right (grain.c:31, 65, 83, 106)
  q.t_impact = gp->impact_time_ma * 1.000e+6 * 365.250 * 24.000 * 3600.000;
  q.t_window_end = (gp->impact_time_ma + 0.45) * 1.000e+6 * 365.250 * 24.000 * 3600.000;
  q.t_start = gp->start_time_ma * 1.000e+6 * 365.250 * 24.000 * 3600.000;
  q.timestep = c->timestep;
  q.xsize = c->P.xsize;
  q.ysize = c->P.ysize;
  DBuf<float> gr;
  DBuf<int> orig;
  ST_OK(gr.alloc(M));
  std::vector<float> slot(c->permuted ? M : 0);
  std::vector<int> o;
  if (c->permuted) {
    o.resize(M);
    HIP_OK(hipMemcpy(o.data(), c->morig.p, sizeof(int) * M, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < M; i++) slot[i] = markgr[o[i]];
  }
  int st = I2E_OK;
  if (hipMemcpyAsync(gr.p, c->permuted ? slot.data() : markgr, sizeof(float) * M, hipMemcpyHostToDevice, c->s) !=
      hipSuccess) {
    set_error("grain: upload of the grain sizes failed");
    st = I2E_ERR_HIP;
  }
  if (!st) {
    const unsigned gm = mgrid_v(M);
    hipLaunchKernelGGL(k_grain, dim3(gm), dim3(256), 0, c->s, q, c->mk, M, gr.p);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(c->permuted ? slot.data() : markgr, gr.p, sizeof(float) * M, hipMemcpyDeviceToHost, c->s) !=
            hipSuccess ||
        hipStreamSynchronize(c->s) != hipSuccess) {
      set_error("grain: kernel or download failed");
      st = I2E_ERR_HIP;
    }
  }
  if (!st && c->permuted)
    for (int64_t i = 0; i < M; i++) markgr[o[i]] = slot[i];
  gr.release();
  return st;
}

int i2e_visual_write(const char *path, const i2e_visual_spec *sp, double timesum, const float *val,
                     const int32_t *rock) {
  if (!path || !sp || (sp->type == 5 ? !rock : !val) || sp->xresol < 2 || sp->yresol < 2) return I2E_ERR_ARG;
  FILE *fl = fopen(path, "wt");
  if (!fl) {
    set_error("visual: cannot open %s", path);
    return I2E_ERR_ARG;
  }
  const long xr = sp->xresol, yr = sp->yresol;
  fprintf(fl, "%e \n %ld %ld \n", timesum / 3.15576e+7, xr, yr);
  for (long m1 = 0; m1 < xr; m1++)
    for (long m2 = 0; m2 < yr; m2++) {
      const long m4 = m1 * yr + m2;
      if (sp->type != 5) {
        fprintf(fl, " % 5.4e", val[m4]);
      } else {
        long m5 = 1, m3;
        for (m3 = m2 + 1; m3 < yr; m3++) {
          if (rock[m4] == rock[m4 + m3 - m2])
            m5++;
          else
            break;
        }
        if (m5 > 3) {
          fprintf(fl, " %d %ld %ld", -2, m5, (long)rock[m4]);
          m2 = m3 - 1;
        } else {
          fprintf(fl, " %ld", (long)rock[m4]);
        }
      }
    }
  fclose(fl);
  return I2E_OK;
}

int i2e_markersave(const char *dir, const i2e_params *p, const i2e_host_state *s, const double *melt) {
  if (!dir || !p || !s || (s->marknum && (!s->markx || !s->marky || !s->markt))) return I2E_ERR_ARG;
  const int64_t M = s->marknum;
  const char *names[4] = {"marker_melt.txt", "marker_comp.txt", "marker_magnet.txt", "marker_magnet_time.txt"};
  for (int f = 0; f < 4; f++) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", dir, names[f]);
    FILE *fl = fopen(path, "wt");
    if (!fl) {
      set_error("markersave: cannot open %s", path);
      return I2E_ERR_ARG;
    }
    // marknum + 1 records, the last one zero (the file layout of loadmart.c:745-779)
    for (int64_t m3 = 0; m3 <= M; m3++) {
      const bool z = m3 == M;
      const float x = z ? 0.0f : s->markx[m3], y = z ? 0.0f : s->marky[m3];
      const double xs = x / p->xsize, ys = y / p->ysize;
      switch (f) {
        case 0: fprintf(fl, "% 5.4e % 5.4e % 5.4e \n", xs, ys, (z || !melt) ? 0.0 : melt[m3]); break;
        case 1: fprintf(fl, "% 5.4e % 5.4e %d \n", xs, ys, z ? 0 : (int)s->markt[m3]); break;
        case 2: fprintf(fl, "% 5.4e % 5.4e %d \n", xs, ys, (z || !s->markmg_old) ? 0 : (int)s->markmg_old[m3]); break;
        default:
          fprintf(fl, "% 5.4e % 5.4e % 5.4e \n", xs, ys, (z || !s->markmg_time) ? 0.0 : s->markmg_time[m3]);
          break;
      }
    }
    fclose(fl);
  }
  return I2E_OK;
}

}  // extern "C"
