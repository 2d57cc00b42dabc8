/*
 * i2e_oracle.c -- CPU FP64 restatement of the I2ELVIS marker-in-cell step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only to check or time against.  The product path never
 * calls it.
 *
 * It follows the reference's serial C code function by function (file:line
 * cited at each function) with the same index maps, operation order for the
 * marker arithmetic (markmart.c) and the same row semantics for the three
 * sparse systems that gausmat4 hands to PARDISO (movemart.c, heatmart.c).
 * Linear systems are returned as 0-based CSR (rows in the order gausmat4
 * appends them, duplicates summed in order of appearance, zero sums
 * dropped -- gausmart.c:320-403) and solved by the Python side.
 *
 * Scope restrictions (each returns I2E_ERR_UNSUPPORTED rather than guessing):
 * intermod = intermod1 = 0, stoksfd = heatfd = 0 (mode.t3c:6011-6036),
 * densimod in {0,1}, adiabyn = 0.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "../include/i2e.h"

#define ABSV(a) ((a) >= 0 ? (a) : -(a))
#define MAXV(a, b) ((a) >= (b) ? (a) : (b))
#define MINV(a, b) ((a) <= (b) ? (a) : (b))

typedef struct orc_csr {
  int64_t n, cap, nnz;
  int32_t *ia; /* n+1 */
  int32_t *ja; /* cap */
  double *a;   /* cap */
  double *b;   /* n   */
} orc_csr;

/* ------------------------------------------------------------------------ */
/* grid helpers                                                              */
/* ------------------------------------------------------------------------ */
typedef struct G {
  const double *gx, *gy;
  int64_t X, Y;
  double xsize, ysize;
} G;

static G mkG(const i2e_params *p, const i2e_host_state *s) {
  G g;
  g.gx = s->gx;
  g.gy = s->gy;
  g.X = p->xnumx;
  g.Y = p->ynumy;
  g.xsize = p->xsize;
  g.ysize = p->ysize;
  return g;
}

/* m1serch / m2serch, markmart.c:1051-1088 */
static int64_t serch(const double *g, int64_t n, double x) {
  int64_t m, m0 = 0, m1 = n - 1;
  do {
    m = (m0 + m1) / 2;
    if (g[m] > x)
      m1 = m;
    else
      m0 = m;
  } while ((m1 - m0) > 1);
  if (m0 > n - 2) m0 = n - 2;
  return m0;
}

int64_t orc_m1serch(const double *gx, int64_t xnumx, double x) { return serch(gx, xnumx, x); }

/* Fornberg weights for two points (fdweight with n=1, movemart.c:1269-1319).
 * m=0: interpolation weights; m=1: first-derivative weights.  The operation
 * order reproduces the recurrence exactly for n=1.                          */
static inline void fd2_interp(double x0, double x1, double xi, double *w0, double *w1) {
  double c4 = x0 - xi, c5, c3;
  c5 = c4;
  c4 = x1 - xi;
  c3 = x1 - x0;
  *w1 = -1.0 * c5 * 1.0 / c3;
  *w0 = c4 * 1.0 / c3;
}
static inline void fd2_deriv(double x0, double x1, double *d0, double *d1) {
  double c3 = x1 - x0;
  *d1 = 1.0 / c3;
  *d0 = -1.0 / c3;
}

/* node-position of sub-grid line m (nodewt, markmart.c:1304-1325):
 * sh=0 basic line, sh=-1 midpoint with previous, sh=+1 with next.          */
static inline double subline(const double *g, int64_t m, int sh) {
  if (sh < 0) return (g[m] + g[m - 1]) / 2.0;
  if (sh > 0) return (g[m] + g[m + 1]) / 2.0;
  return g[m];
}

typedef struct W4 {
  int64_t i0, j0;
  double wx0, wx1, wy0, wy1;
} W4;

static inline void nodewt2(const G *g, int64_t i0, int64_t j0, double x, double y, int ynx, int yny, W4 *w) {
  w->i0 = i0;
  w->j0 = j0;
  fd2_interp(subline(g->gx, i0, ynx), subline(g->gx, i0 + 1, ynx), x, &w->wx0, &w->wx1);
  fd2_interp(subline(g->gy, j0, yny), subline(g->gy, j0 + 1, yny), y, &w->wy0, &w->wy1);
}

/* sum order m1 outer, m2 inner, starting from 0 (markmart.c:1415-1422) */
static inline double isum(const G *g, const W4 *w, const double *f) {
  int64_t m = w->i0 * g->Y + w->j0;
  double e = 0;
  e += (w->wx0 * w->wy0) * f[m];
  e += (w->wx0 * w->wy1) * f[m + 1];
  e += (w->wx1 * w->wy0) * f[m + g->Y];
  e += (w->wx1 * w->wy1) * f[m + g->Y + 1];
  return e;
}
static inline double isum2(const G *g, const W4 *w, const double *f, const double *h) {
  int64_t m = w->i0 * g->Y + w->j0;
  double e = 0;
  e += (w->wx0 * w->wy0) * f[m] * h[m];
  e += (w->wx0 * w->wy1) * f[m + 1] * h[m + 1];
  e += (w->wx1 * w->wy0) * f[m + g->Y] * h[m + g->Y];
  e += (w->wx1 * w->wy1) * f[m + g->Y + 1] * h[m + g->Y + 1];
  return e;
}

static inline void clampxy(const G *g, double *x, double *y) {
  if (*x < 0)
    *x = 0;
  else if (*x > g->xsize)
    *x = g->xsize;
  if (*y < 0)
    *y = 0;
  else if (*y > g->ysize)
    *y = g->ysize;
}

/* stencil origins of the four staggered sub-grids (allinterv 1378-1453,
 * allintere 1458-1543, allintert 1687-1745), intermod = 0                  */
static inline int64_t vx_j0(const G *g, int64_t m20, double y) {
  int64_t j = m20;
  if (y < (g->gy[m20] + g->gy[m20 + 1]) / 2.0) j -= 1;
  if (j < 0) j = 0;
  if (j > g->Y - 3) j = g->Y - 3;
  return j;
}
static inline int64_t vy_i0(const G *g, int64_t m10, double x) {
  int64_t i = m10;
  if (x < (g->gx[m10] + g->gx[m10 + 1]) / 2.0) i -= 1;
  if (i < 0) i = 0;
  if (i > g->X - 3) i = g->X - 3;
  return i;
}
static inline int64_t bn_i0(int64_t m10, int64_t n) { /* exy at basic nodes */
  int64_t i = m10;
  if (i < 1) i = 1;
  if (i > n - 3) i = n - 3;
  return i;
}
static inline int64_t ce_i0(const double *gg, int64_t m10, double x, int64_t n) { /* cell centres */
  int64_t i = m10;
  if (x > (gg[m10] + gg[m10 + 1]) / 2.0) i += 1;
  if (i < 1) i = 1;
  if (i > n - 2) i = n - 2;
  return i;
}

typedef struct Ev { /* the reference's eps[] slots */
  double vx, vy;                        /* eps[11], eps[12] */
  double exy, sxy, exx, sxx, eyy, syy;  /* eps[4..9]        */
  double pr, ga;                        /* eps[10], eps[49] */
  int64_t m10, m20;                     /* wn[0], wn[1]     */
} Ev;

/* allinterv, markmark.c:1378-1453 */
static void allinterv(const G *g, const double *vx, const double *vy, double x, double y, double *ovx, double *ovy) {
  W4 w;
  int64_t m10, m20;
  clampxy(g, &x, &y);
  m10 = serch(g->gx, g->X, x);
  m20 = serch(g->gy, g->Y, y);
  nodewt2(g, m10, vx_j0(g, m20, y), x, y, 0, +1, &w);
  *ovx = isum(g, &w, vx);
  nodewt2(g, vy_i0(g, m10, x), m20, x, y, +1, 0, &w);
  *ovy = isum(g, &w, vy);
}

void orc_allinterv(const i2e_params *p, const i2e_host_state *s, double x, double y, double *out2) {
  G g = mkG(p, s);
  allinterv(&g, s->vx, s->vy, x, y, &out2[0], &out2[1]);
}

/* allintere, markmart.c:1458-1543 */
static void allintere(const G *g, const i2e_host_state *s, double x, double y, Ev *e) {
  W4 w;
  int64_t m10, m20;
  clampxy(g, &x, &y);
  e->m10 = m10 = serch(g->gx, g->X, x);
  e->m20 = m20 = serch(g->gy, g->Y, y);
  nodewt2(g, bn_i0(m10, g->X), bn_i0(m20, g->Y), x, y, 0, 0, &w);
  e->exy = isum(g, &w, s->exy);
  e->sxy = isum(g, &w, s->sxy);
  nodewt2(g, ce_i0(g->gx, m10, x, g->X), ce_i0(g->gy, m20, y, g->Y), x, y, -1, -1, &w);
  e->exx = isum(g, &w, s->exx);
  e->sxx = isum(g, &w, s->sxx);
  e->eyy = isum(g, &w, s->eyy);
  e->syy = isum(g, &w, s->syy);
  e->pr = isum(g, &w, s->pr);
  e->ga = isum(g, &w, s->ga);
}

/* allintere at one point for the tests: out8 = {exy, sxy, exx, sxx, eyy,
 * syy, pr, ga}                                                               */
void orc_allintere(const i2e_params *p, const i2e_host_state *s, double x, double y, double *out8) {
  G g = mkG(p, s);
  Ev e;
  allintere(&g, s, x, y, &e);
  out8[0] = e.exy;
  out8[1] = e.sxy;
  out8[2] = e.exx;
  out8[3] = e.sxx;
  out8[4] = e.eyy;
  out8[5] = e.syy;
  out8[6] = e.pr;
  out8[7] = e.ga;
}

/* ------------------------------------------------------------------------ */
/* rheology, density (markmart.c:1855-2063, 1024-1046)                       */
/* ------------------------------------------------------------------------ */
static double viscalc(const i2e_params *p, double mtk, double mpb, int mm2, double epsin) {
  const i2e_rock *r = &p->rock[mm2];
  double rt = 8.314 * mtk, e1, k1, sduct, nunewt = 0, nupowl = 0, nueff;
  if (mm2 > 20) return r->n0;
  if (r->nu0 != 0) {
    if (r->dh == 0 && r->dv == 0 && (r->ss == 0 || r->mm == 1.0)) nunewt = 1.0 / r->nu0;
    if ((r->dh != 0 || r->dv != 0) && (r->ss == 0 || r->mm == 1.0)) {
      e1 = (r->dh + r->dv * mpb) / rt;
      if (e1 > 150.0) e1 = 150.0;
      if (r->dh < 0 && r->dv > 0) {
        e1 = r->dh * (mtk - r->dv);
        if (e1 < -150.0) e1 = -150.0;
      }
      if (r->dh < 0 && r->dv < 0) {
        e1 = (-r->dh) * (1.0 - (mtk - (-r->dv)) / (-r->dv)) / 8.314 / (-r->dv);
        if (e1 > 150.0) e1 = 150.0;
      }
      nunewt = 1.0 / (r->nu0 * exp(e1));
    }
    if ((r->dh == 0 && r->dv == 0) && r->ss != 0 && r->mm != 1.0) {
      k1 = r->nu0 / pow(r->ss, r->mm - 1.0);
      nunewt = 1.0 / (0.5 * k1);
      sduct = pow(r->nu0 * epsin, 1.0 / r->mm);
      if (epsin > 0) nupowl = 1.0 / (0.5 * sduct / epsin);
    }
    if ((r->dh != 0 || r->dv != 0) && r->ss != 0 && r->mm != 1.0) {
      e1 = (r->dh + r->dv * mpb) / rt;
      if (e1 > 150.0) e1 = 150.0;
      e1 = exp(e1);
      k1 = r->nu0 / pow(r->ss, r->mm - 1.0);
      nunewt = 1.0 / (0.5 * k1 * e1);
      sduct = pow(r->nu0 * e1 * epsin, 1.0 / r->mm);
      if (epsin > 0) nupowl = 1.0 / (0.5 * sduct / epsin);
    }
  }
  nueff = nunewt + nupowl;
  if (nueff <= 0) nueff = 1.0 / r->n1;
  nueff = 1.0 / nueff;
  if (nueff < r->n0) nueff = r->n0;
  if (nueff > r->n1) nueff = r->n1;
  if (nueff < p->nubeg) nueff = p->nubeg;
  if (nueff > p->nuend) nueff = p->nuend;
  return nueff;
}

double orc_viscalc(const i2e_params *p, double mtk, double mpb, int mm2, double epsin) {
  return viscalc(p, mtk, mpb, mm2, epsin);
}

static double dencalc(const i2e_params *p, double mtk, double mpb, double por, int mm2, double *bbout) {
  const i2e_rock *r = &p->rock[mm2];
  *bbout = 0;
  if (p->densimod == 0) return r->ro;
  *bbout = r->bb / (1.0 - r->bb * (mtk - 298.15));
  return r->ro * (1.0 - r->bb * (mtk - 298.15)) * (1.0 + r->aa * (mpb - 1.0) * 1e-3) * (1.000 - por);
}

/* ------------------------------------------------------------------------ */
/* maxvelstep, movemart.c:1225-1262                                          */
/* ------------------------------------------------------------------------ */
int orc_maxvelstep(const i2e_params *p, const i2e_host_state *s, double *timestep) {
  double maxvx = 0, maxvy = 0, ival;
  int64_t m1, m2, m3, X = p->xnumx, Y = p->ynumy;
  double xstpx = p->xsize / (double)(X - 1), ystpy = p->ysize / (double)(Y - 1);
  for (m1 = 0; m1 < X; m1++)
    for (m2 = 0; m2 < Y; m2++) {
      m3 = m1 * Y + m2;
      if (m2 < Y - 1) {
        ival = ABSV(s->vx[m3]);
        maxvx = MAXV(maxvx, ival);
      }
      if (m1 < X - 1) {
        ival = ABSV(s->vy[m3]);
        maxvy = MAXV(maxvy, ival);
      }
    }
  if (maxvx) {
    maxvx = (p->maxxystep * xstpx) / maxvx;
    *timestep = MINV(maxvx, *timestep);
  }
  if (maxvy) {
    maxvy = (p->maxxystep * ystpy) / maxvy;
    *timestep = MINV(maxvy, *timestep);
  }
  return I2E_OK;
}

/* ------------------------------------------------------------------------ */
/* movemark, markmart.c:3-188 (markmod=2 RK4, no immobile markers)           */
/* ------------------------------------------------------------------------ */
static void move_one(const i2e_params *p, const i2e_host_state *s, const G *g, int64_t mm1, double timestep) {
  int mm2 = (int)s->markt[mm1];
  double vx1, vy1, vx2, vy2, vx3, vy3, vx4, vy4, vx0, vy0, mpb, mtk, mnu, x0, y0;
  Ev e;
  if (mm2 >= 100) mm2 -= 100;
  if (!((s->markx[mm1] >= 0 && s->marky[mm1] >= 0 && (double)(s->markx[mm1]) <= g->xsize &&
         (double)(s->marky[mm1]) <= g->ysize) ||
        p->outgrid != 1))
    return;
  if (mm2 >= 0 && mm2 < I2E_MAXTYPES && p->immobile[mm2]) return; /* markmart.c:27 */
  allintere(g, s, (double)(s->markx[mm1]), (double)(s->marky[mm1]), &e);
  mpb = e.pr * 1e-5;
  mtk = (double)(s->markk[mm1]);
  s->markp[mm1] = (float)e.pr;
  if (s->markt[mm1] < 50 && mtk > 0) {
    double eii = pow(0.5 * (e.exx * e.exx + e.eyy * e.eyy) + e.exy * e.exy, 0.5);
    mnu = viscalc(p, mtk, mpb, s->markt[mm1] < 20 ? mm2 : mm2 - 20, eii);
    s->markv[mm1] = (float)mnu;
  }
  x0 = (double)(s->markx[mm1]);
  y0 = (double)(s->marky[mm1]);
  if (p->markmod == 1) { /* Euler, markmart.c:74-79 */
    allinterv(g, s->vx, s->vy, x0, y0, &vx0, &vy0);
  } else { /* RK4, markmart.c:80-97 */
    allinterv(g, s->vx, s->vy, x0, y0, &vx1, &vy1);
    allinterv(g, s->vx, s->vy, x0 + vx1 * timestep / 2.0, y0 + vy1 * timestep / 2.0, &vx2, &vy2);
    allinterv(g, s->vx, s->vy, x0 + vx2 * timestep / 2.0, y0 + vy2 * timestep / 2.0, &vx3, &vy3);
    allinterv(g, s->vx, s->vy, x0 + vx3 * timestep, y0 + vy3 * timestep, &vx4, &vy4);
    vx0 = (vx1 + 2.0 * vx2 + 2.0 * vx3 + vx4) / 6.0;
    vy0 = (vy1 + 2.0 * vy2 + 2.0 * vy3 + vy4) / 6.0;
  }
  if (p->outgrid == 2) {
    if (s->markx[mm1] < 0 || (double)(s->markx[mm1]) > g->xsize) vy0 = 0;
    if (s->marky[mm1] < 0 || (double)(s->marky[mm1]) > g->ysize) vx0 = 0;
  }
  s->markx[mm1] += (float)(timestep * vx0);
  s->marky[mm1] += (float)(timestep * vy0);
  if (s->markx[mm1] < 0 || s->marky[mm1] < 0 || (double)(s->markx[mm1]) > g->xsize ||
      (double)(s->marky[mm1]) > g->ysize)
    s->markk[mm1] = 0;
}

int orc_movemark(const i2e_params *p, i2e_host_state *s) {
  G g = mkG(p, s);
  int64_t mm1;
  for (mm1 = 0; mm1 < s->marknum; mm1++) move_one(p, s, &g, mm1, s->timestep);
  return I2E_OK;
}

/* The flow adjoint's objective (no reference counterpart): mass-weighted
 * mean outward velocity of the mobile iron markers inside the grid, each
 * moving with the velocity move_one would give it (RK4, or Euler for
 * markmod 1; markmart.c:74-97), m = markro of its type.                     */
int orc_sink_rate(const i2e_params *p, const i2e_host_state *s, double *J) {
  G g = mkG(p, s);
  double smv = 0, sm = 0, dt = s->timestep;
  int64_t mm1;
  for (mm1 = 0; mm1 < s->marknum; mm1++) {
    int t = (int)s->markt[mm1], mm2 = t >= 100 ? t - 100 : t;
    double x0 = (double)s->markx[mm1], y0 = (double)s->marky[mm1], vx0, vy0, dx, dy, r;
    if (!(s->markx[mm1] > 0 && s->marky[mm1] > 0 && x0 < g.xsize && y0 < g.ysize && s->markk[mm1] > 0 && t < 50))
      continue;
    if (mm2 < 0 || mm2 >= I2E_MAXTYPES || !p->is_iron[mm2] || p->immobile[mm2]) continue;
    dx = x0 - g.xsize / 2.0;
    dy = y0 - g.ysize / 2.0;
    r = sqrt(dx * dx + dy * dy);
    if (!(r > 0)) continue;
    if (p->markmod == 1) {
      allinterv(&g, s->vx, s->vy, x0, y0, &vx0, &vy0);
    } else {
      double vx1, vy1, vx2, vy2, vx3, vy3, vx4, vy4;
      allinterv(&g, s->vx, s->vy, x0, y0, &vx1, &vy1);
      allinterv(&g, s->vx, s->vy, x0 + vx1 * dt / 2.0, y0 + vy1 * dt / 2.0, &vx2, &vy2);
      allinterv(&g, s->vx, s->vy, x0 + vx2 * dt / 2.0, y0 + vy2 * dt / 2.0, &vx3, &vy3);
      allinterv(&g, s->vx, s->vy, x0 + vx3 * dt, y0 + vy3 * dt, &vx4, &vy4);
      vx0 = (vx1 + 2.0 * vx2 + 2.0 * vx3 + vx4) / 6.0;
      vy0 = (vy1 + 2.0 * vy2 + 2.0 * vy3 + vy4) / 6.0;
    }
    smv += p->rock[mm2].ro * (vx0 * dx / r + vy0 * dy / r);
    sm += p->rock[mm2].ro;
  }
  *J = sm > 0 ? smv / sm : 0.0;
  return I2E_OK;
}

/* ------------------------------------------------------------------------ */
/* ronurecalc, markmart.c:192-1019                                           */
/* ------------------------------------------------------------------------ */
/* Henke et al. cold pressing / hot sintering of silicate porosity,
 * markmart.c:363-468.  The reference retries the Newton solve without bound
 * (goto recalculate); here retries stop after 64 restarts.                  */
static void porosity_update(const i2e_params *p, i2e_host_state *s, int64_t mm1, double mpb, double mtk,
                            double timestep) {
  const double E_act = 85000.000, fact_a = 4.000e-5, grain_orig = 1.000e-6, pivalue = 3.141592654,
               r_gas = 1.9872, p_ref = 0.130;
  float *por = &s->markpor[mm1];
  if (s->markk[mm1] < 700.000) {
    *por = (float)(MINV((double)(*por), (0.420 + 0.460 * pow((pow((mpb / p_ref), 1.720) + 1.000), (-1.000)))));
    *por = (float)(MINV((double)(*por), 0.850));
  }
  if (s->markk[mm1] >= 700.000) {
    double c_fact, cont_r, funct, funct_drv, funct_error, grain_corr = 0, grain_guess, grain_old = 0,
                                                           grain_stress, por_change = 0, por_orig;
    int n_count = 1, no_iterate;
    por_orig = MINV(p->por_init, (0.420 + 0.460 * pow((pow((mpb / p_ref), 1.720) + 1.000), (-1.000))));
    por_orig = MINV(por_orig, 0.850);
    grain_guess = grain_orig;
    for (;;) {
      for (no_iterate = 1; no_iterate < 100; no_iterate = no_iterate + 1) {
        grain_old = grain_guess;
        funct = pow(grain_guess, 3.000) + pow(grain_orig, 3.000) -
                3.000 * pow(grain_guess, 2.000) * grain_orig *
                    pow(((1.000 - por_orig) / (1.000 - (double)(*por))), (1.000 / 3.000)) +
                pow(grain_orig, 3.000) * (1.000 - por_orig) / (1.000 - (double)(*por));
        funct_drv = 3.000 * pow(grain_guess, 2.000) -
                    6.000 * grain_guess * grain_orig *
                        pow(((1.000 - por_orig) / (1.000 - (double)(*por))), (1.000 / 3.000));
        grain_guess = grain_old - (funct / funct_drv);
      }
      funct_error = ABSV(grain_guess - grain_old);
      if (((funct_error > 5.000e-2) || (grain_guess <= grain_orig)) && n_count < 64) {
        n_count = n_count + 1;
        grain_guess = 1.250 * (double)(n_count)*grain_orig;
        continue;
      }
      break;
    }
    if (grain_guess > grain_orig) grain_corr = grain_guess;
    cont_r = sqrt(pow(grain_corr, 2.000) -
                  pow(grain_orig, 2.000) * pow(((1.000 - por_orig) / (1.000 - (double)(*por))), (2.000 / 3.000)));
    c_fact = 2.000 * sqrt(3.000) * (pow(grain_corr, 2.000) - pow(cont_r, 2.000));
    grain_stress = c_fact * mpb / (pivalue * pow(cont_r, 2.000));
    if (*por > 0.100)
      por_change = fact_a * (1.000 - (double)(*por)) * pow(grain_stress, 1.500) / pow((100.000 * grain_corr), 3.000) *
                   exp(-E_act / (r_gas * mtk)) * timestep;
    if (*por <= 0.100)
      por_change = 10.000 * (double)(*por) * fact_a * (1.000 - (double)(*por)) * pow(grain_stress, 1.500) /
                   pow((100.000 * grain_corr), 3.000) * exp(-E_act / (r_gas * mtk)) * timestep;
    if (por_change > 0.000) *por = (float)((double)(*por) - por_change);
  }
}

/* per-node accumulators of ronurecalc (ro0..tk0, sol0/sol1 and their
 * second halves for the viscosity weights), caller-zeroed                   */
typedef struct Acc {
  double *ro0, *et0, *cp0, *kt0, *ht0, *tk0, *nu0, *nd0, *w0, *w1, *wnu, *wnd;
} Acc;

/* conductivity from the soft-turbulence heat flux (markmart.c:520-535) */
static double soft_turb_k(double kt, float dTf, double mro, double bb, double mga, double mcp, double eta_f,
                          double mnu) {
  double heat_flux = 0.089 * pow(kt, (2.000 / 3.000)) * pow((double)ABSV(dTf), (4.000 / 3.000)) *
                     pow((mro * mro * bb * mga * mcp / eta_f), (1.000 / 3.000));
  if (heat_flux <= 0.0001) return kt;
  return pow((heat_flux / 0.089), (3.000 / 2.000)) / (pow((double)dTf, 2.000) * mro) *
         pow((bb * mga * mcp / mnu), (-1.000 / 2.000));
}

static void marker_props_scatter(const i2e_params *p, i2e_host_state *s, const G *g, int64_t mm1, Acc *A) {
  int mm2 = (int)s->markt[mm1];
  double mpb, mtk, mnu, celdx, celdy, mro, mbb, mcp, mkt, mht = 0, radmax, timesumA, xmelt;
  Ev e;
  W4 w;
  if (mm2 >= 100) mm2 -= 100;
  if (!(s->markx[mm1] > 0 && s->marky[mm1] > 0 && (double)(s->markx[mm1]) < g->xsize &&
        (double)(s->marky[mm1]) < g->ysize && s->markk[mm1] > 0 && s->markt[mm1] < 50))
    return;
  radmax = g->xsize / 2.0 * p->radmax_frac;
  allintere(g, s, (double)(s->markx[mm1]), (double)(s->marky[mm1]), &e);
  mpb = e.pr * 1.0000e-5;
  mtk = (double)(s->markk[mm1]);
  if (mm2 < 2) mtk = s->markk[mm1] = (float)(p->tmp_ambient);
  mnu = s->markv[mm1];
  if (mnu <= 0) {
    double eii = pow(0.5 * (e.exx * e.exx + e.eyy * e.eyy) + e.exy * e.exy, 0.5);
    mnu = viscalc(p, mtk, mpb, s->markt[mm1] < 20 ? mm2 : mm2 - 20, eii);
  }
  celdx = s->markx[mm1] - g->xsize / 2.0;
  celdy = s->marky[mm1] - g->ysize / 2.0;
  celdy = pow(celdx * celdx + celdy * celdy, 0.5);
  if (celdy > radmax) mnu = p->eta_outer;
  if (mm2 == 0 || mm2 >= 7) s->markpor[mm1] = 0.000;
  if (s->markt[mm1] != 0)
    if (s->marktmax[mm1] < s->markk[mm1]) s->marktmax[mm1] = s->markk[mm1];
  if (p->porosity_on && p->porous[mm2] && (s->markpor[mm1] > 0.000 && s->markpor[mm1] <= 0.850))
    porosity_update(p, s, mm1, mpb, mtk, s->timestep);
  if (s->markpor[mm1] < 0.000) s->markpor[mm1] = 0.000;
  if (s->markpor[mm1] > 1.000) s->markpor[mm1] = 1.000;
  mro = dencalc(p, mtk, mpb, (double)(s->markpor[mm1]), mm2, &mbb);
  mcp = p->rock[mm2].cp;
  mkt = (p->rock[mm2].kt + p->rock[mm2].kf / (mtk + 77.0)) * exp(p->rock[mm2].kp * mpb);
  /* melting closure in place of tdbasecalc (markmart.c:2178-2220) */
  xmelt = 0;
  if (p->melt_on && p->meltable[mm2 < 20 ? mm2 : mm2 - 20]) {
    double xm = (mtk - p->melt_ts) / (p->melt_tl - p->melt_ts), xp, xq;
    int tnew;
    if (xm < 0.0) xm = 0.0;
    if (xm > 1.0) xm = 1.0;
    xmelt = xm;
    tnew = s->markt[mm1];
    if (xm > 0 && tnew < 20) tnew += 20;
    if (xm <= 0 && tnew > 20) tnew -= 20;
    s->markt[mm1] = (int8_t)tnew;
    if (tnew > 20) {
      const i2e_rock *rm = &p->rock[tnew];
      mro = mro * (1.0 - xm) + xm * rm->ro * (1.0 - rm->bb * (mtk - 298.15)) * (1.0 + rm->aa * (mpb - 1.0) * 1e-3);
      if (xm > p->melt_switch) {
        double nueff = p->melt_eta;
        if (nueff < p->nubeg) nueff = p->nubeg;
        if (nueff > p->nuend) nueff = p->nuend;
        mnu = nueff;
      }
      mkt = ((rm->kt + rm->kf / (mtk + 77.0)) * exp(rm->kp * mpb)) * xm +
            ((p->rock[tnew - 20].kt + p->rock[tnew - 20].kf / (mtk + 77.0)) * exp(p->rock[tnew - 20].kp * mpb)) *
                (1.0 - xm);
      if (xm > 0 && xm < 1.0) {
        xp = (mtk + 1.0 - p->melt_ts) / (p->melt_tl - p->melt_ts);
        xq = (mtk - 1.0 - p->melt_ts) / (p->melt_tl - p->melt_ts);
        if (xp < 0.0) xp = 0.0;
        if (xp > 1.0) xp = 1.0;
        if (xq < 0.0) xq = 0.0;
        if (xq > 1.0) xq = 1.0;
        mcp += (p->melt_hlat * xp - p->melt_hlat * xq) / (2.0 * 1.0);
      }
    }
  }
  /* soft-turbulence conductivity, markmart.c:503-610.  The melt fractions
   * are the marker's own (the reference reads eps[21]/eps[50] as the last
   * melt-fraction call left them); dT against marker 0's temperature. */
  if (p->soft_turb && s->timesum > p->core_form_time) {
    const float dTf = s->markk[mm1] - s->markk[0];
    const i2e_rock *r = &p->rock[mm2];
    if (mm2 == 25 || mm2 == 26) {
      double eta_f = -1.0;
      if (xmelt >= 0.4000 && xmelt <= 0.6000) eta_f = r->n0 * exp((-172.690) * (xmelt - 0.400));
      if (xmelt > 0.6000) eta_f = 100.0000;
      if (eta_f > 0) mkt = soft_turb_k(r->kt, dTf, mro, r->bb, e.ga, mcp, eta_f, mnu);
    }
    if (mm2 == 7 || mm2 == 8 || mm2 == 9 || mm2 == 10 || mm2 == 17 || mm2 == 18 || mm2 == 19) {
      double xfe = 0;
      if (p->fe_tl > 0) {
        xfe = (mtk - p->fe_ts) / (p->fe_tl - p->fe_ts);
        if (xfe < 0.000) xfe = 0.0000;
        if (xfe > 1.000) xfe = 1.0000;
      }
      mkt = soft_turb_k(r->kt, dTf, mro, r->bb, e.ga, mcp, xfe < 1.0000 ? 1.0000e12 : 0.010, mnu);
    }
  }
  /* porosity-dependent conductivity, markmart.c:612-638 */
  if (p->porous[mm2]) {
    double por_ref1 = 0.080, por_ref2 = 0.167, a_fact = -1.200, kt0 = p->rock[mm2].kt, pr0 = s->markpor[mm1];
    if (s->markpor[mm1] < 0.200) mkt = kt0 * exp(-pr0 / por_ref1);
    if (s->markpor[mm1] >= 0.200 && s->markpor[mm1] <= 0.400)
      mkt = pow((pow((kt0 * exp(-pr0 / por_ref1)), 4.000) + pow((kt0 * exp(a_fact - (pr0 / por_ref2))), 4.000)), 0.250);
    if (s->markpor[mm1] > 0.400) mkt = kt0 * exp(a_fact - (pr0 / por_ref2));
  }
  if (mkt <= 0.0010) mkt = 0.0010;
  if (mkt >= p->k_cutoff) mkt = p->k_cutoff;
  /* magnetisation of solid silicates, markmart.c:651-744 (a state the
   * rules do not assign keeps the marker's own time; the reference keeps
   * the previous marker's) */
  if (p->magnet && s->markmg_old && s->markmg_time) {
    int mag = 3, old = s->markmg_old[mm1], sil = (mm2 == 5 || mm2 == 6);
    double mag_time = s->markmg_time[mm1];
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
          if (old <= 1 || old == 3) mag_time = s->timesum / (3600.000 * 24.000 * 365.250 * 1.000e+6);
        } else if (mtk > 851.16 && mtk <= 948.16) {
          mag = 1;
          if (old == 0 || old == 3) mag_time = s->timesum / (3600.000 * 24.000 * 365.250 * 1.000e+6);
        } else {
          mag = 0;
          mag_time = -1.000;
        }
      } else {
        mag = 3;
        mag_time = -1.000;
      }
    }
    s->markmg_old[mm1] = (int8_t)mag;
    s->markmg_time[mm1] = mag_time;
  }
  /* radiogenic heating, markmart.c:747-831 */
  timesumA = s->timesum / (365.250 * 24.000 * 3600.000);
  if (mm2 == 0) mht = 0.0000e-10;
  if (p->is_silicate[mm2]) {
    double init_al = 1.89716e+23 * p->al2627_init * 4.999e-13 / (3.264e+13);
    double mht_al = init_al * exp(-timesumA * 9.6673e-7) * p->h0_scale, ll = 0;
    if (p->longlived != 0)
      ll = 0.40689 * 1.430e-11 * exp(-timesumA * 4.9867e-10) + 2.990e-12 * exp(-timesumA * 9.8458e-10) +
           1.600e-12 * exp(-timesumA * 1.5541e-10) + 1.000e-12 * exp(-timesumA * 4.9511e-11);
    mht = (mht_al + ll) * mro;
  } else if (p->is_iron[mm2]) {
    double init_fe = 1.971e+24 * p->fe6056_init * 4.3451e-13 / (1.183e+14);
    double mht_fe = init_fe * exp(-timesumA * 2.6660e-7), ka = 0;
    if (p->longlived != 0) ka = 0.090625 * 1.430e-11 * exp(-timesumA * 4.9867e-10);
    mht = (mht_fe + ka) * mro;
  }
  /* scatter, markmart.c:840-911 (intermod1 = 0, viscmod = 0) */
  nodewt2(g, e.m10, e.m20, (double)(s->markx[mm1]), (double)(s->marky[mm1]), 0, 0, &w);
  {
    double ddx = 1.0 - ABSV(w.wx0 - 0.5), ddy = 1.0 - ABSV(w.wy0 - 0.5), swt2;
    int64_t m3 = (e.m10 + 1) * g->Y + e.m20 + 1;
    if (ddx > 0.5 && ddy > 0.5) {
      swt2 = 1.0 * 4.0 * (ddx - 0.5) * (ddy - 0.5);
      A->nd0[m3] += mnu * swt2;
      A->wnd[m3] += swt2;
    }
  }
  {
    int a, b;
    for (a = 0; a < 2; a++)
      for (b = 0; b < 2; b++) {
        int64_t m3 = (e.m10 + a) * g->Y + e.m20 + b;
        double ddx = a ? w.wx1 : w.wx0, ddy = b ? w.wy1 : w.wy0;
        double swt = 1.0 * ddx * ddy, swt2;
        if (ddx > 0.5 && ddy > 0.5) {
          swt2 = 1.0 * 4.0 * (ddx - 0.5) * (ddy - 0.5);
          A->nu0[m3] += mnu * swt2;
          A->wnu[m3] += swt2;
        }
        A->ro0[m3] += mro * swt;
        A->et0[m3] += mbb * swt;
        A->cp0[m3] += mcp * swt;
        A->kt0[m3] += mkt * swt;
        A->ht0[m3] += mht * swt;
        A->w0[m3] += swt;
        if (!p->immobile[mm2]) { /* markmart.c:905 */
          A->tk0[m3] += mtk * swt;
          A->w1[m3] += swt;
        }
      }
  }
}

/* marker loop + node normalisation + viscosity-contrast limit
 * (markmart.c:259-1009).  tkrecalc follows separately.  out2 = {numin,numax} */
int orc_ronurecalc(const i2e_params *p, i2e_host_state *s, double *out2) {
  G g = mkG(p, s);
  int64_t N = p->xnumx * p->ynumy, m3, mm1;
  double numin = 0, numax = 0;
  Acc A;
  double *buf;
  if (p->densimod > 1) return I2E_ERR_UNSUPPORTED;
  buf = (double *)calloc((size_t)N * 12, sizeof(double));
  if (!buf) return I2E_ERR_CAPACITY;
  A.ro0 = buf;
  A.et0 = buf + N;
  A.cp0 = buf + 2 * N;
  A.kt0 = buf + 3 * N;
  A.ht0 = buf + 4 * N;
  A.tk0 = buf + 5 * N;
  A.nu0 = buf + 6 * N;
  A.nd0 = buf + 7 * N;
  A.w0 = buf + 8 * N;
  A.w1 = buf + 9 * N;
  A.wnu = buf + 10 * N;
  A.wnd = buf + 11 * N;
  for (mm1 = 0; mm1 < s->marknum; mm1++) marker_props_scatter(p, s, &g, mm1, &A);
  for (m3 = 0; m3 < N; m3++) {
    if (A.wnu[m3]) {
      A.nu0[m3] /= A.wnu[m3];
      if (A.nu0[m3] < p->nubeg) A.nu0[m3] = p->nubeg;
      if (A.nu0[m3] > p->nuend) A.nu0[m3] = p->nuend;
      if (numin == 0 || A.nu0[m3] < numin) numin = A.nu0[m3];
      if (numax == 0 || A.nu0[m3] > numax) numax = A.nu0[m3];
      s->nu[m3] = A.nu0[m3];
    }
    if (A.wnd[m3]) {
      A.nd0[m3] /= A.wnd[m3];
      if (A.nd0[m3] < p->nubeg) A.nd0[m3] = p->nubeg;
      if (A.nd0[m3] > p->nuend) A.nd0[m3] = p->nuend;
      if (numin == 0 || A.nd0[m3] < numin) numin = A.nd0[m3];
      if (numax == 0 || A.nd0[m3] > numax) numax = A.nd0[m3];
      s->nd[m3] = A.nd0[m3];
    }
    if (A.w0[m3]) {
      s->ro[m3] = A.ro0[m3] / A.w0[m3];
      s->et[m3] = A.et0[m3] / A.w0[m3];
      s->cp[m3] = A.cp0[m3] / A.w0[m3];
      s->kt[m3] = A.kt0[m3] / A.w0[m3];
      s->ht[m3] = A.ht0[m3] / A.w0[m3];
      if (A.w1[m3]) s->tk[m3] = A.tk0[m3] / A.w1[m3];
    }
  }
  if (p->nucontr >= 1.0)
    numax = numin * p->nucontr;
  else
    numin = numax * p->nucontr;
  for (m3 = 0; m3 < N; m3++) {
    if (s->nu[m3] < numin) s->nu[m3] = numin;
    if (s->nu[m3] > numax) s->nu[m3] = numax;
    if (s->nd[m3] < numin) s->nd[m3] = numin;
    if (s->nd[m3] > numax) s->nd[m3] = numax;
  }
  if (out2) {
    out2[0] = numin;
    out2[1] = numax;
  }
  free(buf);
  return I2E_OK;
}

/* per-cell marker-type histogram (cell of a marker = (m1serch, m2serch) of
 * its clamped position, the cell every scatter above uses)                  */
int orc_type_counts(const i2e_params *p, const i2e_host_state *s, int ntypes, int32_t *cnt) {
  G g = mkG(p, s);
  int64_t mm1, C = (p->xnumx - 1) * (p->ynumy - 1);
  memset(cnt, 0, sizeof(int32_t) * (size_t)(C * ntypes));
  for (mm1 = 0; mm1 < s->marknum; mm1++) {
    double x = s->markx[mm1], y = s->marky[mm1];
    int t = s->markt[mm1];
    if (!(s->markx[mm1] > 0 && s->marky[mm1] > 0 && x < g.xsize && y < g.ysize)) continue;
    if (t < 0 || t >= ntypes) continue;
    cnt[(serch(g.gx, g.X, x) * (p->ynumy - 1) + serch(g.gy, g.Y, y)) * ntypes + t]++;
  }
  return I2E_OK;
}

/* hydration and body statistics, impact.c:3022-3228: the node pass
 * (:3064-3076), the rheology switch and the classification of every silicate
 * marker by its current type and peak temperature (:3078-3206), the fractions
 * (:3208-3223), in the reference's order                                    */
int orc_hydration(const i2e_params *p, i2e_host_state *s, double t_sol, double t_liq, int hydrate,
                  i2e_body_stats *o) {
  int64_t m1, mm1, N = p->xnumx * p->ynumy;
  int64_t count_toohot = 0, count_body_nodes = 0, count_tot = 0, count_sol = 0, count_liq = 0, count_hydrous = 0,
          count_primitive = 0, count_n2co2 = 0, count_cocl = 0, count_h2o = 0, count_phyllo1 = 0, count_phyllo2 = 0,
          count_phyllo3 = 0, count_phyllo4 = 0, count_perco = 0, count_melt1 = 0, count_melt2 = 0;
  double maxtk = p->tmp_ambient, meantk = 0.0, t_max_body = p->tmp_ambient, t_mean_body = 0.0;
  for (m1 = 0; m1 < N; m1++) {
    maxtk = MAXV(maxtk, s->tk[m1]);
    if (s->tk[m1] != p->tmp_ambient) {
      meantk += s->tk[m1];
      count_body_nodes += 1;
    }
  }
  meantk = meantk / (double)(count_body_nodes);
  if (hydrate) {
    for (mm1 = 0; mm1 < s->marknum; mm1++) {
      if (s->markt[mm1] == 5 && s->marktmax[mm1] >= 273.15 && s->marktmax[mm1] < 1223.15) s->markt[mm1] = 6;
      if (s->markt[mm1] == 6 && s->marktmax[mm1] >= 1223.15) s->markt[mm1] = 5;
      if (s->markt[mm1] == 5 || s->markt[mm1] == 6) count_sol += 1;
      if (s->markt[mm1] == 25 || s->markt[mm1] == 26) count_liq += 1;
      if (s->markt[mm1] == 5 || s->markt[mm1] == 6 || s->markt[mm1] == 25 || s->markt[mm1] == 26) {
        if (s->marktmax[mm1] < 273.15) count_primitive += 1;
        if (s->marktmax[mm1] >= 273.15 && s->marktmax[mm1] < 1223.15) count_hydrous += 1;
        if (s->marktmax[mm1] < 1073.15) count_n2co2 += 1;
        if (s->marktmax[mm1] < (0.10 * (t_liq - t_sol) + t_sol)) count_cocl += 1;
        if (s->marktmax[mm1] < 1223.15) count_h2o += 1;
        if (s->marktmax[mm1] >= 273.15 && s->marktmax[mm1] < 773.15) count_phyllo1 += 1;
        if (s->marktmax[mm1] >= 773.15 && s->marktmax[mm1] < 873.15) count_phyllo2 += 1;
        if (s->marktmax[mm1] >= 873.15 && s->marktmax[mm1] < 973.15) count_phyllo3 += 1;
        if (s->marktmax[mm1] >= 973.15 && s->marktmax[mm1] < 1223.15) count_phyllo4 += 1;
        if (s->marktmax[mm1] >= 1273.15) count_perco += 1;
        if (s->marktmax[mm1] >= t_sol) count_melt1 += 1;
        if (s->marktmax[mm1] >= (0.40 * (t_liq - t_sol) + t_sol)) count_melt2 += 1;
        count_tot += 1;
        if (s->markk[mm1] > t_max_body) t_max_body = s->markk[mm1];
        t_mean_body += s->markk[mm1];
        if (s->markk[mm1] / maxtk > 1.1) count_toohot += 1;
      }
    }
  }
  o->sol_frac = (double)(count_sol) / (double)(count_tot);
  o->liq_frac = (double)(count_liq) / (double)(count_tot);
  o->hydrous_frac = (double)(count_hydrous) / (double)(count_tot);
  o->primitive_frac = (double)(count_primitive) / (double)(count_tot);
  o->n2co2_frac = (double)(count_n2co2) / (double)(count_tot);
  o->cocl_frac = (double)(count_cocl) / (double)(count_tot);
  o->h2o_frac = (double)(count_h2o) / (double)(count_tot);
  o->phyllo1_frac = (double)(count_phyllo1) / (double)(count_tot);
  o->phyllo2_frac = (double)(count_phyllo2) / (double)(count_tot);
  o->phyllo3_frac = (double)(count_phyllo3) / (double)(count_tot);
  o->phyllo4_frac = (double)(count_phyllo4) / (double)(count_tot);
  o->perco_frac = (double)(count_perco) / (double)(count_tot);
  o->melt1_frac = (double)(count_melt1) / (double)(count_tot);
  o->melt2_frac = (double)(count_melt2) / (double)(count_tot);
  o->t_mean_body = t_mean_body / (double)(count_tot);
  o->maxtk = maxtk;
  o->meantk = meantk;
  o->t_max_body = t_max_body;
  o->count_toohot = count_toohot;
  o->count_tot = count_tot;
  o->count_body_nodes = count_body_nodes;
  return I2E_OK;
}

/* ------------------------------------------------------------------------ */
/* gausmat4 row builder, gausmart.c:320-403                                  */
/* ------------------------------------------------------------------------ */
typedef struct Row {
  int n;
  int64_t col[64];
  double v[64];
  double rhs;
} Row;

static inline void radd(Row *r, int64_t col, double v) {
  r->col[r->n] = col;
  r->v[r->n] = v;
  r->n++;
}

static int emit(orc_csr *A, int64_t row, const Row *r) {
  int idx[64], i, j, k = 0, n = 0;
  int64_t cols[64];
  double sums[64];
  /* entries with wi != 0 in order of appearance, then columns ascending with
   * per-column sums accumulated in appearance order */
  for (i = 0; i < r->n; i++)
    if (r->v[i]) idx[n++] = i;
  for (i = 0; i < n; i++) {
    int64_t c = r->col[idx[i]];
    for (j = 0; j < k; j++)
      if (cols[j] == c) break;
    if (j == k) {
      cols[k] = c;
      sums[k] = 0.0;
      k++;
    }
    sums[j] += r->v[idx[i]];
  }
  /* insertion sort by column */
  for (i = 1; i < k; i++) {
    int64_t c = cols[i];
    double s = sums[i];
    j = i - 1;
    while (j >= 0 && cols[j] > c) {
      cols[j + 1] = cols[j];
      sums[j + 1] = sums[j];
      j--;
    }
    cols[j + 1] = c;
    sums[j + 1] = s;
  }
  A->b[row] = r->rhs;
  A->ia[row] = (int32_t)A->nnz;
  for (i = 0; i < k; i++)
    if (sums[i]) {
      if (A->nnz >= A->cap) return I2E_ERR_CAPACITY;
      A->ja[A->nnz] = (int32_t)cols[i];
      A->a[A->nnz] = sums[i];
      A->nnz++;
    }
  A->ia[row + 1] = (int32_t)A->nnz;
  return I2E_OK;
}

/* generic BC row x_i - sum k x_nbr = const (x/y/pbonderr, tbonderr),
 * neighbour numbers 1-based in bondn, offset subtracted for T rows          */
static void bcrow(const i2e_host_state *s, int64_t self, int64_t slot, int64_t off, Row *r) {
  int64_t b = s->bondm[slot];
  int n1;
  r->n = 0;
  r->rhs = s->bondv[b * 4 + 0];
  radd(r, self, 1.0);
  for (n1 = 0; n1 < 3; n1++)
    if (s->bondn[b * 3 + n1]) radd(r, s->bondn[b * 3 + n1] - 1 - off, -s->bondv[b * 4 + n1 + 1]);
}

static inline void rscale(Row *r, double k) {
  int i;
  r->rhs *= k;
  for (i = 0; i < r->n; i++) r->v[i] *= k;
}

/* ------------------------------------------------------------------------ */
/* gravity potential, movemart.c:622-785                                     */
/* ------------------------------------------------------------------------ */
int orc_gravity_build(const i2e_params *p, i2e_host_state *s, orc_csr *A) {
  int64_t X = p->xnumx, Y = p->ynumy, m1, m2, mcmax;
  const double *gx = s->gx, *gy = s->gy;
  double gamma = 6.67384e-11, pival = 2.0 * asin(1.0), rosum = 0, ronum = 0, dx, dy, ival, ival1, mukoef,
         GYKOEF = p->gykoef, GXKOEF = s->gxkoef;
  double xstpx = p->xsize / (double)(X - 1), ystpy = p->ysize / (double)(Y - 1);
  Row r;
  int st;
  for (m1 = 1; m1 < X; m1++)
    for (m2 = 1; m2 < Y; m2++) {
      mcmax = m1 * Y + m2;
      dx = (gx[m1 - 1] + gx[m1] - p->xsize) / 2.0;
      dy = (gy[m2 - 1] + gy[m2] - p->ysize) / 2.0;
      ival = pow(dx * dx + dy * dy, 0.5);
      if (ival < GYKOEF) {
        ival1 = (gx[m1] - gx[m1 - 1]) * (gy[m2] - gy[m2 - 1]);
        rosum += ival1 * (s->ro[mcmax - 1 - Y] + s->ro[mcmax - Y] + s->ro[mcmax - 1] + s->ro[mcmax]) / 4.0;
        ronum += ival1;
      }
    }
  if (ronum > 0) {
    rosum /= ronum;
    GXKOEF = 4.0 / 3.0 * pival * GYKOEF * rosum * gamma;
  }
  s->gxkoef = GXKOEF;
  mukoef = 2.0 / xstpx / xstpx + 2 / ystpy / ystpy;
  A->nnz = 0;
  A->n = X * Y;
  for (m1 = 0; m1 < X; m1++)
    for (m2 = 0; m2 < Y; m2++) {
      mcmax = m1 * Y + m2;
      r.n = 0;
      if (m1 > 0 && m2 > 0) {
        dx = (gx[m1 - 1] + gx[m1] - p->xsize) / 2.0;
        dy = (gy[m2 - 1] + gy[m2] - p->ysize) / 2.0;
        ival = pow(dx * dx + dy * dy, 0.5);
        if (m1 == 1 || m1 == X - 1 || m2 == 1 || m2 == Y - 1) {
          r.rhs = -GXKOEF * GYKOEF * GYKOEF / ival * mukoef;
          radd(&r, mcmax, 1.0 * mukoef);
        } else {
          double dx02, dx24, dy12, dy23;
          r.rhs = 4.0 / 1.5 * pival * gamma *
                  (s->ro[mcmax - 1 - Y] + s->ro[mcmax - Y] + s->ro[mcmax - 1] + s->ro[mcmax]) / 4.0;
          dx02 = (gx[m1] - gx[m1 - 2]) / 2.0;
          dx24 = (gx[m1 + 1] - gx[m1 - 1]) / 2.0;
          radd(&r, mcmax - Y, 2.0 / dx02 / (dx02 + dx24));
          radd(&r, mcmax, -2.0 / dx02 / (dx02 + dx24));
          radd(&r, mcmax + Y, 2.0 / dx24 / (dx02 + dx24));
          radd(&r, mcmax, -2.0 / dx24 / (dx02 + dx24));
          dy12 = (gy[m2] - gy[m2 - 2]) / 2.0;
          dy23 = (gy[m2 + 1] - gy[m2 - 1]) / 2.0;
          radd(&r, mcmax - 1, 2.0 / dy12 / (dy12 + dy23));
          radd(&r, mcmax, -2.0 / dy12 / (dy12 + dy23));
          radd(&r, mcmax + 1, 2.0 / dy23 / (dy12 + dy23));
          radd(&r, mcmax, -2.0 / dy23 / (dy12 + dy23));
          if (ival > GYKOEF) {
            r.rhs -= GXKOEF * GYKOEF * GYKOEF / ival * mukoef * 1e+18;
            radd(&r, mcmax, 1.0 * mukoef * 1e+18);
          }
        }
      } else {
        r.rhs = 0;
        radd(&r, mcmax, 1.0 * mukoef);
      }
      if ((st = emit(A, mcmax, &r))) return st;
    }
  return I2E_OK;
}

/* gp reload + |g| at basic nodes, movemart.c:757-783 */
void orc_gravity_post(const i2e_params *p, i2e_host_state *s, const double *x) {
  int64_t X = p->xnumx, Y = p->ynumy, m1, m2, mcmax, N = X * Y;
  const double *gx = s->gx, *gy = s->gy;
  double *gp = s->gp, ival, ival1;
  for (m1 = 0; m1 < N; m1++) gp[m1] = x[m1];
  for (m1 = 1; m1 < X - 1; m1++)
    for (m2 = 1; m2 < Y - 1; m2++) {
      mcmax = m1 * Y + m2;
      ival = (gp[mcmax + Y] + gp[mcmax + Y + 1] - gp[mcmax] - gp[mcmax + 1]) / (gx[m1 + 1] - gx[m1 - 1]);
      ival1 = (gp[mcmax + 1] + gp[mcmax + Y + 1] - gp[mcmax] - gp[mcmax + 1]) / (gy[m2 + 1] - gy[m2 - 1]);
      s->ga[mcmax] = pow(ival * ival + ival1 * ival1, 0.500);
    }
}

/* ------------------------------------------------------------------------ */
/* Stokes + continuity rows, movemart.c:4-279, 287-426, 789-967, 1326-1649   */
/* ------------------------------------------------------------------------ */
/* coefficients of sigma_xx at cell (m1,m2) times ynval (sxxcalc 1326-1412) */
static void sxx_row(const i2e_params *p, const i2e_host_state *s, int64_t m1, int64_t m2, double ynval, Row *r) {
  int64_t Y = p->ynumy;
  double nueff = s->nd[m1 * Y + m2], d0, d1;
  fd2_deriv(s->gx[m1 - 1], s->gx[m1], &d0, &d1);
  radd(r, ((m1 - 1) * Y + (m2 - 1)) * 3 + 1, ynval * 2.0 * nueff * d0);
  radd(r, (m1 * Y + (m2 - 1)) * 3 + 1, ynval * 2.0 * nueff * d1);
}
static void syy_row(const i2e_params *p, const i2e_host_state *s, int64_t m1, int64_t m2, double ynval, Row *r) {
  int64_t Y = p->ynumy;
  double nueff = s->nd[m1 * Y + m2], d0, d1;
  fd2_deriv(s->gy[m2 - 1], s->gy[m2], &d0, &d1);
  radd(r, ((m1 - 1) * Y + (m2 - 1)) * 3 + 2, ynval * 2.0 * nueff * d0);
  radd(r, ((m1 - 1) * Y + m2) * 3 + 2, ynval * 2.0 * nueff * d1);
}
/* sxycalc 1515-1649 including its clamping of the stencil at the edges */
static void sxy_row(const i2e_params *p, const i2e_host_state *s, int64_t m1, int64_t m2, double ynval, Row *r) {
  int64_t X = p->xnumx, Y = p->ynumy, m2min, m2max, m1min, m1max;
  double nueff = s->nu[m1 * Y + m2], d0 = 0, d1 = 0;
  m2min = m2 - 1;
  if (m2min < 0) m2min = 0;
  m2max = m2;
  if (m2max > Y - 2) m2max = Y - 2;
  if (m2max > m2min) fd2_deriv(subline(s->gy, m2min, 1), subline(s->gy, m2max, 1), &d0, &d1);
  if (m2max > m2min) {
    radd(r, (m1 * Y + m2min) * 3 + 1, ynval * 2.0 * nueff * d0 / 2.0);
    radd(r, (m1 * Y + m2max) * 3 + 1, ynval * 2.0 * nueff * d1 / 2.0);
  } else {
    radd(r, (m1 * Y + m2min) * 3 + 1, ynval * 2.0 * nueff * 0.0 / 2.0);
  }
  m1min = m1 - 1;
  if (m1min < 0) m1min = 0;
  m1max = m1;
  if (m1max > X - 2) m1max = X - 2;
  d0 = d1 = 0;
  if (m1max > m1min) {
    fd2_deriv(subline(s->gx, m1min, 1), subline(s->gx, m1max, 1), &d0, &d1);
    radd(r, (m1min * Y + m2) * 3 + 2, ynval * 2.0 * nueff * d0 / 2.0);
    radd(r, (m1max * Y + m2) * 3 + 2, ynval * 2.0 * nueff * d1 / 2.0);
  } else {
    radd(r, (m1min * Y + m2) * 3 + 2, ynval * 2.0 * nueff * 0.0 / 2.0);
  }
}

static void xstokes_row(const i2e_params *p, const i2e_host_state *s, int64_t m1, int64_t m2, Row *r) {
  int64_t Y = p->ynumy, v0 = m1 * Y + m2, v1 = v0 + 1, v2 = v0 + Y, v3 = v2 + 1;
  const double *gx = s->gx, *gy = s->gy, *ro = s->ro;
  double xkf = (gx[m1 + 1] - gx[m1 - 1]) / 2.0, ykf = gy[m2 + 1] - gy[m2];
  double GX = -(s->gp[v3] - s->gp[v1]) / xkf, dt = s->timestep;
  r->n = 0;
  r->rhs = -GX * (ro[v0] + ro[v1]) / 2.0;
  radd(r, v1 * 3 + 0, +1.0 / xkf);
  radd(r, v3 * 3 + 0, -1.0 / xkf);
  sxx_row(p, s, m1, m2 + 1, -1.0 / xkf, r);
  sxx_row(p, s, m1 + 1, m2 + 1, 1.0 / xkf, r);
  sxy_row(p, s, m1, m2, -1.0 / ykf, r);
  sxy_row(p, s, m1, m2 + 1, 1.0 / ykf, r);
  radd(r, v0 * 3 + 1, -dt / 2.0 * GX * (ro[v2] + ro[v3] - ro[v0 - Y] - ro[v1 - Y]) / (gx[m1 + 1] - gx[m1 - 1]) / 2.0);
  radd(r, v0 * 3 + 2,
       -dt / 2.0 * GX * (gx[m1] - gx[m1 - 1]) / (gx[m1 + 1] - gx[m1 - 1]) / 2.0 * (ro[v1] - ro[v0]) / (gy[m2 + 1] - gy[m2]));
  radd(r, v1 * 3 + 2,
       -dt / 2.0 * GX * (gx[m1] - gx[m1 - 1]) / (gx[m1 + 1] - gx[m1 - 1]) / 2.0 * (ro[v1] - ro[v0]) / (gy[m2 + 1] - gy[m2]));
  radd(r, (v0 - Y) * 3 + 2,
       -dt / 2.0 * GX * (gx[m1 + 1] - gx[m1]) / (gx[m1 + 1] - gx[m1 - 1]) / 2.0 * (ro[v1] - ro[v0]) / (gy[m2 + 1] - gy[m2]));
  radd(r, (v1 - Y) * 3 + 2,
       -dt / 2.0 * GX * (gx[m1 + 1] - gx[m1]) / (gx[m1 + 1] - gx[m1 - 1]) / 2.0 * (ro[v1] - ro[v0]) / (gy[m2 + 1] - gy[m2]));
}

static void ystokes_row(const i2e_params *p, const i2e_host_state *s, int64_t m1, int64_t m2, Row *r) {
  int64_t Y = p->ynumy, v0 = m1 * Y + m2, v1 = v0 + 1, v2 = v0 + Y, v3 = v2 + 1;
  const double *gx = s->gx, *gy = s->gy, *ro = s->ro;
  double xkf = gx[m1 + 1] - gx[m1], ykf = (gy[m2 + 1] - gy[m2 - 1]) / 2.0;
  double GY = -(s->gp[v3] - s->gp[v2]) / ykf, dt = s->timestep;
  r->n = 0;
  r->rhs = -GY * (ro[v0] + ro[v2]) / 2.0;
  radd(r, v2 * 3 + 0, +1.0 / ykf);
  radd(r, v3 * 3 + 0, -1.0 / ykf);
  syy_row(p, s, m1 + 1, m2, -1.0 / ykf, r);
  syy_row(p, s, m1 + 1, m2 + 1, 1.0 / ykf, r);
  sxy_row(p, s, m1, m2, -1.0 / xkf, r);
  sxy_row(p, s, m1 + 1, m2, 1.0 / xkf, r);
  radd(r, v0 * 3 + 2, -dt / 2.0 * GY * (ro[v1] + ro[v3] - ro[v0 - 1] - ro[v2 - 1]) / (gy[m2 + 1] - gy[m2 - 1]) / 2.0);
  radd(r, v0 * 3 + 1,
       -dt / 2.0 * GY * (gy[m2] - gy[m2 - 1]) / (gy[m2 + 1] - gy[m2 - 1]) / 2.0 * (ro[v2] - ro[v0]) / (gx[m1 + 1] - gx[m1]));
  radd(r, v2 * 3 + 1,
       -dt / 2.0 * GY * (gy[m2] - gy[m2 - 1]) / (gy[m2 + 1] - gy[m2 - 1]) / 2.0 * (ro[v2] - ro[v0]) / (gx[m1 + 1] - gx[m1]));
  radd(r, (v0 - 1) * 3 + 1,
       -dt / 2.0 * GY * (gy[m2 + 1] - gy[m2]) / (gy[m2 + 1] - gy[m2 - 1]) / 2.0 * (ro[v2] - ro[v0]) / (gx[m1 + 1] - gx[m1]));
  radd(r, (v2 - 1) * 3 + 1,
       -dt / 2.0 * GY * (gy[m2 + 1] - gy[m2]) / (gy[m2 + 1] - gy[m2 - 1]) / 2.0 * (ro[v2] - ro[v0]) / (gx[m1 + 1] - gx[m1]));
}

/* conterr assembly part, movemart.c:287-426 */
static void cont_row(const i2e_params *p, const i2e_host_state *s, int64_t m1, int64_t m2, Row *r) {
  int64_t Y = p->ynumy;
  double d0, d1;
  r->n = 0;
  r->rhs = 0;
  fd2_deriv(s->gx[m1 - 1], s->gx[m1], &d0, &d1);
  radd(r, ((m1 - 1) * Y + (m2 - 1)) * 3 + 1, d0);
  radd(r, (m1 * Y + (m2 - 1)) * 3 + 1, d1);
  fd2_deriv(s->gy[m2 - 1], s->gy[m2], &d0, &d1);
  radd(r, ((m1 - 1) * Y + (m2 - 1)) * 3 + 2, d0);
  radd(r, ((m1 - 1) * Y + m2) * 3 + 2, d1);
}

/* continuity-type velocity BC of unknown U (movemart.c:863-879): the
 * continuity row (conterr) of node (m1 + dm1, m2 + dm2), the shift read from
 * the first neighbour's unknown; the reference exits on an inconsistent
 * table, this returns I2E_ERR_ARG                                           */
static int cont_bc_row(const i2e_params *p, const i2e_host_state *s, int64_t m1, int64_t m2, int64_t U, int comp,
                       Row *r) {
  int64_t X = p->xnumx, Y = p->ynumy, b = s->bondm[U], u0 = s->bondn[3 * b], m3, dm1 = 0, dm2 = 0;
  if (u0 <= 0) return I2E_ERR_ARG;
  m3 = (u0 - 1 - U) / 3;
  if (m3 >= Y) dm1 = 1;
  if (m3 == 1 || m3 > Y) dm2 = 1;
  if (m3 < 0 || m3 > Y + 1 || !(m1 + dm1) || !(m2 + dm2) || m1 + dm1 > X - 1 || m2 + dm2 > Y - 1 ||
      u0 - 1 - comp < 0 || u0 - 1 - comp >= 3 * X * Y || !s->bondm[u0 - 1 - comp])
    return I2E_ERR_ARG;
  cont_row(p, s, m1 + dm1, m2 + dm2, r);
  return I2E_OK;
}

int orc_stokes_build(const i2e_params *p, const i2e_host_state *s, orc_csr *A) {
  int64_t X = p->xnumx, Y = p->ynumy, m1, m2, mcmax;
  double mukoef = p->nubeg * p->strmin;
  Row r;
  int st;
  A->nnz = 0;
  A->n = 3 * X * Y;
  for (m1 = 0; m1 < X; m1++)
    for (m2 = 0; m2 < Y; m2++) {
      mcmax = (m1 * Y + m2) * 3;
      /* P / continuity */
      if (m1 && m2) {
        if (!s->bondm[mcmax + 0])
          cont_row(p, s, m1, m2, &r);
        else
          bcrow(s, mcmax + 0, mcmax + 0, 0, &r);
        rscale(&r, mukoef);
      } else {
        r.n = 0;
        r.rhs = 0;
        radd(&r, mcmax + 0, mukoef);
      }
      if ((st = emit(A, mcmax + 0, &r))) return st;
      /* Vx */
      if (m2 < Y - 1) {
        if (!s->bondm[mcmax + 1])
          xstokes_row(p, s, m1, m2, &r);
        else {
          if (s->bondv[s->bondm[mcmax + 1] * 4 + 1] > 1e+30) {
            if ((st = cont_bc_row(p, s, m1, m2, mcmax + 1, 1, &r))) return st;
          } else {
            bcrow(s, mcmax + 1, mcmax + 1, 0, &r);
          }
          rscale(&r, mukoef);
        }
      } else {
        r.n = 0;
        r.rhs = 0;
        radd(&r, mcmax + 1, mukoef);
      }
      if ((st = emit(A, mcmax + 1, &r))) return st;
      /* Vy */
      if (m1 < X - 1) {
        if (!s->bondm[mcmax + 2])
          ystokes_row(p, s, m1, m2, &r);
        else {
          if (s->bondv[s->bondm[mcmax + 2] * 4 + 1] > 1e+30) {
            if ((st = cont_bc_row(p, s, m1, m2, mcmax + 2, 2, &r))) return st;
          } else {
            bcrow(s, mcmax + 2, mcmax + 2, 0, &r);
          }
          rscale(&r, mukoef);
        }
      } else {
        r.n = 0;
        r.rhs = 0;
        radd(&r, mcmax + 2, mukoef);
      }
      if ((st = emit(A, mcmax + 2, &r))) return st;
    }
  return I2E_OK;
}

/* solution reload + strain rate / stress recompute, movemart.c:979-1045 */
void orc_stokes_post(const i2e_params *p, i2e_host_state *s, const double *x) {
  int64_t X = p->xnumx, Y = p->ynumy, m1, m2, m;
  for (m1 = 0; m1 < X; m1++)
    for (m2 = 0; m2 < Y; m2++) {
      m = m1 * Y + m2;
      if (m1 && m2) s->pr[m] = x[3 * m + 0];
      if (m2 < Y - 1) s->vx[m] = x[3 * m + 1];
      if (m1 < X - 1) s->vy[m] = x[3 * m + 2];
    }
  for (m1 = 1; m1 < X; m1++)
    for (m2 = 1; m2 < Y; m2++) {
      double d0, d1, e;
      m = m1 * Y + m2;
      fd2_deriv(s->gx[m1 - 1], s->gx[m1], &d0, &d1);
      e = 0;
      e += s->vx[(m1 - 1) * Y + (m2 - 1)] * d0;
      e += s->vx[m1 * Y + (m2 - 1)] * d1;
      s->exx[m] = e;
      s->sxx[m] = e * (2.0 * s->nd[m]);
      fd2_deriv(s->gy[m2 - 1], s->gy[m2], &d0, &d1);
      e = 0;
      e += s->vy[(m1 - 1) * Y + (m2 - 1)] * d0;
      e += s->vy[(m1 - 1) * Y + m2] * d1;
      s->eyy[m] = e;
      s->syy[m] = e * (2.0 * s->nd[m]);
      if (m1 < X - 1 && m2 < Y - 1) {
        e = 0;
        fd2_deriv(subline(s->gy, m2 - 1, 1), subline(s->gy, m2, 1), &d0, &d1);
        e += s->vx[m1 * Y + m2 - 1] * d0 / 2.0;
        e += s->vx[m1 * Y + m2] * d1 / 2.0;
        fd2_deriv(subline(s->gx, m1 - 1, 1), subline(s->gx, m1, 1), &d0, &d1);
        e += s->vy[(m1 - 1) * Y + m2] * d0 / 2.0;
        e += s->vy[m1 * Y + m2] * d1 / 2.0;
        s->exy[m] = e;
        s->sxy[m] = e * (2.0 * s->nu[m]);
      }
    }
}

/* Residual diagnostics of viterate (movemart.c:1049-1176): RMS of the
 * x/y-Stokes errors in their error form (xstokserr/ystokserr with ynerr=1,
 * movemart.c:51-81, 186-222: stresses from the post-solve fields, divided by
 * the mean nodal viscosity), of the continuity errors (conterr, ynerr=1,
 * :354-372) and of the velocity BC rows (x/ybonderr, :442-457), normalised
 * by the velocity range and the smallest grid step.  out[14] = stokes,
 * cont, bond, stoknum, contnum, bondnum, vxmin, vxmax, vymin, vymax,
 * prmin, prmax, numin, numax.                                               */
static double xstokes_errv(const i2e_host_state *s, int64_t Y, int64_t m1, int64_t m2, double *nue, double *pmax,
                           double *pmin) {
  const double *gx = s->gx, *gy = s->gy;
  double xkf = (gx[m1 + 1] - gx[m1 - 1]) / 2.0, ykf = gy[m2 + 1] - gy[m2];
  int64_t v0 = m1 * Y + m2, v1 = v0 + 1, v2 = v0 + Y, v3 = v2 + 1;
  double gxk = -(s->gp[v3] - s->gp[v1]) / xkf;
  double rightx = -gxk * (s->ro[v0] + s->ro[v1]) / 2.0, leftx, nueff;
  leftx = (s->sxx[v3] - s->sxx[v1]) / xkf;
  leftx += (s->sxy[v1] - s->sxy[v0]) / ykf;
  leftx -= (s->pr[v3] - s->pr[v1]) / xkf;
  leftx -= s->timestep / 2.0 * gxk *
           (s->vx[v0] * (s->ro[v2] + s->ro[v3] - s->ro[v0 - Y] - s->ro[v1 - Y]) / (gx[m1 + 1] - gx[m1 - 1]) / 2.0 +
            ((s->vy[v0] + s->vy[v1]) * (gx[m1] - gx[m1 - 1]) + (s->vy[v0 - Y] + s->vy[v1 - Y]) * (gx[m1 + 1] - gx[m1])) /
                (gx[m1 + 1] - gx[m1 - 1]) / 2.0 * (s->ro[v1] - s->ro[v0]) / (gy[m2 + 1] - gy[m2]));
  nueff = (s->nu[v0] + s->nu[v1]) / 2.0;
  *nue = nueff;
  *pmax = MAXV(s->pr[v1], s->pr[v3]);
  *pmin = MINV(s->pr[v1], s->pr[v3]);
  return (leftx - rightx) / nueff;
}
static double ystokes_errv(const i2e_host_state *s, int64_t Y, int64_t m1, int64_t m2, double *nue, double *pmax,
                           double *pmin) {
  const double *gx = s->gx, *gy = s->gy;
  double xkf = gx[m1 + 1] - gx[m1], ykf = (gy[m2 + 1] - gy[m2 - 1]) / 2.0;
  int64_t v0 = m1 * Y + m2, v1 = v0 + 1, v2 = v0 + Y, v3 = v2 + 1;
  double gyk = -(s->gp[v3] - s->gp[v2]) / ykf;
  double righty = -gyk * (s->ro[v0] + s->ro[v2]) / 2.0, lefty, nueff;
  lefty = (s->syy[v3] - s->syy[v2]) / ykf;
  lefty += (s->sxy[v2] - s->sxy[v0]) / xkf;
  lefty -= (s->pr[v3] - s->pr[v2]) / ykf;
  lefty -= s->timestep / 2.0 * gyk *
           (s->vy[v0] * (s->ro[v1] + s->ro[v3] - s->ro[v0 - 1] - s->ro[v2 - 1]) / (gy[m2 + 1] - gy[m2 - 1]) / 2.0 +
            ((s->vx[v0] + s->vx[v2]) * (gy[m2] - gy[m2 - 1]) + (s->vx[v0 - 1] + s->vx[v2 - 1]) * (gy[m2 + 1] - gy[m2])) /
                (gy[m2 + 1] - gy[m2 - 1]) / 2.0 * (s->ro[v2] - s->ro[v0]) / (gx[m1 + 1] - gx[m1]));
  nueff = (s->nu[v0] + s->nu[v2]) / 2.0;
  *nue = nueff;
  *pmax = MAXV(s->pr[v2], s->pr[v3]);
  *pmin = MINV(s->pr[v2], s->pr[v3]);
  return (lefty - righty) / nueff;
}
static double cont_errv(const i2e_host_state *s, int64_t Y, int64_t m1, int64_t m2) {
  double d0, d1, divv = 0;
  fd2_deriv(s->gx[m1 - 1], s->gx[m1], &d0, &d1);
  divv += s->vx[(m1 - 1) * Y + (m2 - 1)] * d0;
  divv += s->vx[m1 * Y + (m2 - 1)] * d1;
  fd2_deriv(s->gy[m2 - 1], s->gy[m2], &d0, &d1);
  divv += s->vy[(m1 - 1) * Y + (m2 - 1)] * d0;
  divv += s->vy[(m1 - 1) * Y + m2] * d1;
  return divv;
}
/* value of Stokes unknown u (0-based, 3m+{0,1,2}) from the node fields */
static double sol_at(const i2e_host_state *s, int64_t u) {
  int64_t m = u / 3;
  return u % 3 == 0 ? s->pr[m] : (u % 3 == 1 ? s->vx[m] : s->vy[m]);
}
static double bond_errv(const i2e_host_state *s, int64_t u) {
  int64_t b = s->bondm[u], n1;
  double leftx = sol_at(s, u) - s->bondv[4 * b + 0];
  for (n1 = 0; n1 < 3; n1++)
    if (s->bondn[3 * b + n1]) leftx -= s->bondv[4 * b + n1 + 1] * sol_at(s, s->bondn[3 * b + n1] - 1);
  return leftx;
}
int orc_stokes_residuals(const i2e_params *p, const i2e_host_state *s, double *out) {
  int64_t X = p->xnumx, Y = p->ynumy, m1, m2;
  double minvx = 1e200, maxvx = -1e200, minvy = 1e200, maxvy = -1e200;
  double stoksum = 0, stoknum = 0, contsum = 0, contnum = 0, bondsum = 0, bondnum = 0;
  double minnu = 1e200, maxnu = -1e200, minpr = 1e200, maxpr = -1e200, ival, nue, pmx, pmn;
  double xstpx = p->xsize / (double)(X - 1), ystpy = p->ysize / (double)(Y - 1), mindx;
  for (m1 = 0; m1 < X; m1++)
    for (m2 = 0; m2 < Y; m2++) {
      int64_t m3 = m1 * Y + m2;
      if (m2 < Y - 1) {
        minvx = MINV(minvx, s->vx[m3]);
        maxvx = MAXV(maxvx, s->vx[m3]);
      }
      if (m1 < X - 1) {
        minvy = MINV(minvy, s->vy[m3]);
        maxvy = MAXV(maxvy, s->vy[m3]);
      }
    }
  out[6] = minvx;
  out[7] = maxvx;
  out[8] = minvy;
  out[9] = maxvy;
  maxvx -= minvx;
  maxvy -= minvy;
  maxvx = MAXV(maxvx, maxvy);
  mindx = MINV(xstpx, ystpy);
  for (m1 = 0; m1 < X; m1++)
    for (m2 = 0; m2 < Y; m2++) {
      int64_t mc = (m1 * Y + m2) * 3;
      if (m1 && m2) {
        ival = cont_errv(s, Y, m1, m2);
        contsum += ival * ival;
        contnum += 1.0;
      }
      if (m2 < Y - 1) {
        if (!s->bondm[mc + 1]) {
          if (m1 < 1 || m1 > X - 2) return -1; /* the error form needs interior x-Stokes rows */
          ival = xstokes_errv(s, Y, m1, m2, &nue, &pmx, &pmn);
          stoksum += ival * ival;
          stoknum += 1.0;
          minnu = MINV(minnu, nue);
          maxnu = MAXV(maxnu, nue);
          maxpr = MAXV(maxpr, pmx);
          minpr = MINV(minpr, pmn);
        } else if (s->bondv[4 * s->bondm[mc + 1] + 1] < 1e+30) {
          ival = bond_errv(s, mc + 1);
          bondsum += ival * ival;
          bondnum += 1.0;
        }
      }
      if (m1 < X - 1) {
        if (!s->bondm[mc + 2]) {
          if (m2 < 1 || m2 > Y - 2) return -1;
          ival = ystokes_errv(s, Y, m1, m2, &nue, &pmx, &pmn);
          stoksum += ival * ival;
          stoknum += 1.0;
          minnu = MINV(minnu, nue);
          maxnu = MAXV(maxnu, nue);
          maxpr = MAXV(maxpr, pmx);
          minpr = MINV(minpr, pmn);
        } else if (s->bondv[4 * s->bondm[mc + 2] + 1] < 1e+30) {
          ival = bond_errv(s, mc + 2);
          bondsum += ival * ival;
          bondnum += 1.0;
        }
      }
    }
  out[0] = pow(stoksum / stoknum, 0.5) / (maxvx / mindx / mindx);
  out[1] = pow(contsum / contnum, 0.5) / (maxvx / mindx);
  out[2] = pow(bondsum / bondnum, 0.5) / maxvx;
  out[3] = stoknum;
  out[4] = contnum;
  out[5] = bondnum;
  out[10] = minpr;
  out[11] = maxpr;
  out[12] = minnu;
  out[13] = maxnu;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* heat, heatmart.c:6-155 (heaterr), 162-312 (qx/qycalc), 320-695 (titerate) */
/* ------------------------------------------------------------------------ */
static double qx_val(const i2e_params *p, const double *tk, const double *kt, const double *gx, int64_t m1, int64_t m2) {
  int64_t Y = p->ynumy;
  double kteff = (kt[m1 * Y + m2] + kt[(m1 + 1) * Y + m2]) / 2.0, d0, d1, q = 0;
  fd2_deriv(gx[m1], gx[m1 + 1], &d0, &d1);
  q += tk[m1 * Y + m2] * d0 * kteff;
  q += tk[(m1 + 1) * Y + m2] * d1 * kteff;
  return q;
}
static double qy_val(const i2e_params *p, const double *tk, const double *kt, const double *gy, int64_t m1, int64_t m2) {
  int64_t Y = p->ynumy;
  double kteff = (kt[m1 * Y + m2] + kt[m1 * Y + m2 + 1]) / 2.0, d0, d1, q = 0;
  fd2_deriv(gy[m2], gy[m2 + 1], &d0, &d1);
  q += tk[m1 * Y + m2] * d0 * kteff;
  q += tk[m1 * Y + m2 + 1] * d1 * kteff;
  return q;
}

/* heaterr(m1,m2,0): source term (saved to sol1) and total rate             */
static double heat_source(const i2e_params *p, const i2e_host_state *s, int64_t m1, int64_t m2, double *rate) {
  G g = mkG(p, s);
  int64_t Y = p->ynumy, v4 = m1 * Y + m2;
  double ival = 0, xkf = (s->gx[m1 + 1] - s->gx[m1 - 1]) / 2.0, ykf = (s->gy[m2 + 1] - s->gy[m2 - 1]) / 2.0;
  ival += s->ht[v4] / s->cp[v4] / s->ro[v4];
  if (p->frictyn > 0 && s->timesum > p->frictyn && s->ro[v4] > 1e+3) {
    /* allinters at the node, markmart.c:1548-1682 (eps[13..15]) */
    double x = s->gx[m1], y = s->gy[m2], e13, e14, e15, epsval;
    int64_t m10, m20;
    W4 w;
    clampxy(&g, &x, &y);
    m10 = serch(g.gx, g.X, x);
    m20 = serch(g.gy, g.Y, y);
    nodewt2(&g, bn_i0(m10, g.X), bn_i0(m20, g.Y), x, y, 0, 0, &w);
    e13 = isum2(&g, &w, s->exy, s->sxy);
    nodewt2(&g, ce_i0(g.gx, m10, x, g.X), ce_i0(g.gy, m20, y, g.Y), x, y, -1, -1, &w);
    e14 = isum2(&g, &w, s->exx, s->sxx);
    e15 = isum2(&g, &w, s->eyy, s->syy);
    epsval = 2.0 * e13 + e14 + e15;
    ival += epsval / s->cp[v4] / s->ro[v4];
  }
  if (rate)
    *rate = ival + ((qx_val(p, s->tk, s->kt, s->gx, m1, m2) - qx_val(p, s->tk, s->kt, s->gx, m1 - 1, m2)) / xkf +
                    (qy_val(p, s->tk, s->kt, s->gy, m1, m2) - qy_val(p, s->tk, s->kt, s->gy, m1, m2 - 1)) / ykf) /
                       s->cp[v4] / s->ro[v4];
  return ival;
}

/* titerate (heatmart.c:333-387): sol1 source terms of every interior non-BC
 * node and the global timestep shrink.  Returns tktimestep (0 when the
 * reference returns early), timestep updated in place for heatmod == 0.     */
int orc_heat_prepare(const i2e_params *p, i2e_host_state *s, double *sol1, double *tktimestep_out) {
  int64_t X = p->xnumx, Y = p->ynumy, N = X * Y, m1, m2;
  double maxtkrate = 0, tktimestep = 0, rate;
  if (p->adiabyn > 0) return I2E_ERR_UNSUPPORTED;
  memset(sol1, 0, sizeof(double) * (size_t)N);
  for (m1 = 0; m1 < X; m1++)
    for (m2 = 0; m2 < Y; m2++) {
      int64_t m = m1 * Y + m2;
      if (s->bondm[3 * N + m]) continue;
      if (m1 == 0 || m2 == 0 || m1 == X - 1 || m2 == Y - 1) return I2E_ERR_UNSUPPORTED;
      sol1[m] = heat_source(p, s, m1, m2, &rate);
      rate = ABSV(rate);
      maxtkrate = MAXV(maxtkrate, rate);
    }
  if (maxtkrate) tktimestep = p->maxtkstep / maxtkrate;
  *tktimestep_out = tktimestep;
  if (s->timestep == 0 || tktimestep == 0) return I2E_OK;
  if (p->heatmod == 0) s->timestep = MINV(s->timestep, tktimestep);
  return I2E_OK;
}

int orc_heat_build(const i2e_params *p, const i2e_host_state *s, const double *sol1, orc_csr *A) {
  int64_t X = p->xnumx, Y = p->ynumy, N = X * Y, m1, m2;
  double dt = s->timestep;
  Row r;
  int st;
  A->nnz = 0;
  A->n = N;
  for (m1 = 0; m1 < X; m1++)
    for (m2 = 0; m2 < Y; m2++) {
      int64_t v4 = m1 * Y + m2;
      if (!s->bondm[3 * N + v4]) {
        double xkf = (s->gx[m1 + 1] - s->gx[m1 - 1]) / 2.0, ykf = (s->gy[m2 + 1] - s->gy[m2 - 1]) / 2.0;
        double cpv = s->cp[v4], rov = s->ro[v4], yn, k, d0, d1;
        r.n = 0;
        r.rhs = sol1[v4] + s->tk[v4] / dt;
        radd(&r, v4, 1.0 / dt);
        /* qxcalc(m1-1,m2, 1/xkf/cp/ro) */
        yn = 1.0 / xkf / cpv / rov;
        k = (s->kt[(m1 - 1) * Y + m2] + s->kt[m1 * Y + m2]) / 2.0;
        fd2_deriv(s->gx[m1 - 1], s->gx[m1], &d0, &d1);
        radd(&r, (m1 - 1) * Y + m2, yn * d0 * k);
        radd(&r, m1 * Y + m2, yn * d1 * k);
        /* qxcalc(m1,m2,-1/xkf/cp/ro) */
        yn = -1.0 / xkf / cpv / rov;
        k = (s->kt[m1 * Y + m2] + s->kt[(m1 + 1) * Y + m2]) / 2.0;
        fd2_deriv(s->gx[m1], s->gx[m1 + 1], &d0, &d1);
        radd(&r, m1 * Y + m2, yn * d0 * k);
        radd(&r, (m1 + 1) * Y + m2, yn * d1 * k);
        /* qycalc(m1,m2-1, 1/ykf/cp/ro) */
        yn = 1.0 / ykf / cpv / rov;
        k = (s->kt[m1 * Y + m2 - 1] + s->kt[m1 * Y + m2]) / 2.0;
        fd2_deriv(s->gy[m2 - 1], s->gy[m2], &d0, &d1);
        radd(&r, m1 * Y + m2 - 1, yn * d0 * k);
        radd(&r, m1 * Y + m2, yn * d1 * k);
        /* qycalc(m1,m2,-1/ykf/cp/ro) */
        yn = -1.0 / ykf / cpv / rov;
        k = (s->kt[m1 * Y + m2] + s->kt[m1 * Y + m2 + 1]) / 2.0;
        fd2_deriv(s->gy[m2], s->gy[m2 + 1], &d0, &d1);
        radd(&r, m1 * Y + m2, yn * d0 * k);
        radd(&r, m1 * Y + m2 + 1, yn * d1 * k);
      } else {
        bcrow(s, v4, 3 * N + v4, 3 * N, &r);
      }
      if ((st = emit(A, v4, &r))) return st;
    }
  return I2E_OK;
}

/* tkrecalc system (heatmart.c:758-826): identity rows for free nodes, BC rows
 * otherwise; b[] of free rows is the current tk                              */
int orc_tbc_build(const i2e_params *p, const i2e_host_state *s, orc_csr *A) {
  int64_t N = p->xnumx * p->ynumy, m;
  Row r;
  int st;
  A->nnz = 0;
  A->n = N;
  for (m = 0; m < N; m++) {
    if (!s->bondm[3 * N + m]) {
      r.n = 0;
      r.rhs = s->tk[m];
      radd(&r, m, 1.0);
    } else
      bcrow(s, m, 3 * N + m, 3 * N, &r);
    if ((st = emit(A, m, &r))) return st;
  }
  return I2E_OK;
}

/* marker temperature update after the implicit solve, heatmart.c:515-689.
 * tk = new node T, tk2 = node T before the solve.                            */
int orc_heat_markers(const i2e_params *p, i2e_host_state *s, const double *tk2) {
  G g = mkG(p, s);
  int64_t N = p->xnumx * p->ynumy, m3;
  double *sol0 = (double *)calloc((size_t)N * 2, sizeof(double)), *sol1;
  double dt = s->timestep;
  if (!sol0) return I2E_ERR_CAPACITY;
  sol1 = sol0 + N;
  for (m3 = 0; m3 < s->marknum; m3++) {
    double x = s->markx[m3], y = s->marky[m3], e2, e3, dtk, dtk1, mkt, mcp, mro, celdx, celdy;
    int64_t m10, m20;
    W4 w;
    if (!(s->markx[m3] > 0 && s->marky[m3] > 0 && x < g.xsize && y < g.ysize && s->markt[m3] < 50)) continue;
    m10 = serch(g.gx, g.X, x);
    m20 = serch(g.gy, g.Y, y);
    nodewt2(&g, m10, m20, x, y, 0, 0, &w);
    e2 = isum(&g, &w, s->tk);
    e3 = isum(&g, &w, tk2);
    if (s->markk[m3] <= 0) s->markk[m3] = (float)(e3);
    if (p->heatdif) {
      dtk = e3 - (double)(s->markk[m3]);
      celdx = g.gx[m10 + 1] - g.gx[m10];
      celdy = g.gy[m20 + 1] - g.gy[m20];
      mkt = isum(&g, &w, s->kt);
      mcp = isum(&g, &w, s->cp);
      mro = isum(&g, &w, s->ro);
      dtk1 = -p->heatdif * mkt / mro / mcp * 2.0 * (1.0 / celdx / celdx + 1.0 / celdy / celdy) * dt;
      if (dtk1 < -150.0) dtk1 = -150.0;
      dtk1 = dtk * (1.0 - exp(dtk1));
      s->markk[m3] += (float)(dtk1);
      {
        int a, b;
        for (a = 0; a < 2; a++)
          for (b = 0; b < 2; b++) {
            int64_t m4 = (m10 + a) * g.Y + m20 + b;
            double ival = (a ? w.wx1 : w.wx0) * (b ? w.wy1 : w.wy0) * 1.0;
            sol0[m4] += ival;
            sol1[m4] += dtk1 * ival;
          }
      }
    }
    s->markk[m3] += (float)(e2 - e3);
  }
  if (p->heatdif) {
    for (m3 = 0; m3 < N; m3++)
      if (sol0[m3]) sol1[m3] /= sol0[m3];
    for (m3 = 0; m3 < s->marknum; m3++) {
      double x = s->markx[m3], y = s->marky[m3];
      int64_t m10, m20;
      W4 w;
      int a, b;
      if (!(s->markx[m3] > 0 && s->marky[m3] > 0 && x < g.xsize && y < g.ysize && s->markt[m3] < 50)) continue;
      m10 = serch(g.gx, g.X, x);
      m20 = serch(g.gy, g.Y, y);
      nodewt2(&g, m10, m20, x, y, 0, 0, &w);
      for (a = 0; a < 2; a++)
        for (b = 0; b < 2; b++) {
          int64_t m4 = (m10 + a) * g.Y + m20 + b;
          double ival = (a ? w.wx1 : w.wx0) * (b ? w.wy1 : w.wy0);
          if (sol1[m4]) s->markk[m3] -= (float)(sol1[m4] * ival);
        }
    }
  }
  free(sol0);
  return I2E_OK;
}

/* CSR sparse matrix-vector product, used by the tests to compare the HIP
 * operators with the assembled rows                                          */
void orc_csr_matvec(const orc_csr *A, const double *x, double *y) {
  int64_t i;
  for (i = 0; i < A->n; i++) {
    double s = 0;
    int32_t k;
    for (k = A->ia[i]; k < A->ia[i + 1]; k++) s += A->a[k] * x[A->ja[k]];
    y[i] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* post-processing: the visualisation grid of i2jmart's saver               */
/* ------------------------------------------------------------------------ */
/* allinter1, loadjmart.c:1327-1520 (intermod 0): the values saver() reads */
typedef struct E1 {
  double t, nu, ro, kt, ht;            /* eps[2], eps[16..19]: basic-node cell */
  double exy, sxy, exysxy, esp;        /* eps[4], eps[5], eps[13], eps[27]     */
  double exx, sxx, eyy, syy, pr, exxsxx, eyysyy; /* eps[6..10], eps[14], [15] */
} E1;
static void allinter1(const G *g, const i2e_host_state *s, const double *esp, double x, double y, E1 *e) {
  W4 w;
  int64_t m10, m20;
  if (x < 0) x = 0; else if (x > g->xsize) x = g->xsize;
  if (y < 0) y = 0; else if (y > g->ysize) y = g->ysize;
  m10 = serch(g->gx, g->X, x);
  m20 = serch(g->gy, g->Y, y);
  nodewt2(g, m10, m20, x, y, 0, 0, &w);
  e->t = isum(g, &w, s->tk);
  e->nu = isum(g, &w, s->nu);
  e->ro = isum(g, &w, s->ro);
  e->kt = isum(g, &w, s->kt);
  e->ht = isum(g, &w, s->ht);
  nodewt2(g, bn_i0(m10, g->X), bn_i0(m20, g->Y), x, y, 0, 0, &w);
  e->exy = isum(g, &w, s->exy);
  e->sxy = isum(g, &w, s->sxy);
  e->exysxy = isum2(g, &w, s->exy, s->sxy);
  e->esp = isum(g, &w, esp);
  nodewt2(g, ce_i0(g->gx, m10, x, g->X), ce_i0(g->gy, m20, y, g->Y), x, y, -1, -1, &w);
  e->exx = isum(g, &w, s->exx);
  e->sxx = isum(g, &w, s->sxx);
  e->eyy = isum(g, &w, s->eyy);
  e->syy = isum(g, &w, s->syy);
  e->pr = isum(g, &w, s->pr);
  e->exxsxx = isum2(g, &w, s->exx, s->sxx);
  e->eyysyy = isum2(g, &w, s->eyy, s->syy);
}

/* spncalc, loadjmart.c:1523-1617 (stoksfd 0): the rotation rate at an
 * interior basic node, (dvy/dx - dvx/dy) / 2 with two-point weights        */
static double spncalc(const G *g, const i2e_host_state *s, int64_t m1, int64_t m2) {
  const int64_t Y = g->Y;
  int64_t m2min = m2 - 1 < 0 ? 0 : m2 - 1, m2max = m2 > Y - 2 ? Y - 2 : m2;
  int64_t m1min = m1 - 1 < 0 ? 0 : m1 - 1, m1max = m1 > g->X - 2 ? g->X - 2 : m1;
  double d0, d1, a = 0, b = 0;
  fd2_deriv((g->gy[m2min] + g->gy[m2min + 1]) / 2.0, (g->gy[m2max] + g->gy[m2max + 1]) / 2.0, &d0, &d1);
  a += s->vx[m1 * Y + m2min] * d0 / 2.0;
  a += s->vx[m1 * Y + m2max] * d1 / 2.0;
  fd2_deriv((g->gx[m1min] + g->gx[m1min + 1]) / 2.0, (g->gx[m1max] + g->gx[m1max + 1]) / 2.0, &d0, &d1);
  b += s->vy[m1min * Y + m2] * d0 / 2.0;
  b += s->vy[m1max * Y + m2] * d1 / 2.0;
  return b - a;
}

/* saver(), loadjmart.c:619-1197, for one output and one x-chunk: val (float,
 * xresol*yresol) for spec type != 5, rock (int32) for type 5; melt (marknum,
 * may be NULL) as markersave() reads it.  The markers of s are modified as
 * saver() modifies them (markk < 1 reset, marktmax).                        */
int orc_visual(const i2e_params *p, i2e_host_state *s, const i2e_visual_spec *sp, float *val, int32_t *rock,
               double *melt) {
  const G g = mkG(p, s);
  const int type = sp->type;
  const int64_t M = s->marknum, X = g.X, Y = g.Y;
  int64_t xresol = sp->xresol, yresol = sp->yresol, nresol = xresol * yresol, markstp, m1, m2, m3, m4;
  double xb = sp->xbeg, xe = sp->xend, yb = sp->ybeg, ye = sp->yend;
  double xminx, xmaxx, xresx, yminy, ymaxy, yresy, xminx0, yminy0, ival = 0, ea, na;
  float *val0;
  int64_t *lin0;
  double *esp;
  if (type == 3 || (type >= 9 && type <= 11) || type < 1 || type > 12 || p->densimod == 2) return I2E_ERR_UNSUPPORTED;
  if (xresol < 2 || yresol < 2 || nresol * 2 > 40000000) return I2E_ERR_ARG;
  if (xb < 0) xb = 0;
  if (xe > 1.0) xe = 1.0;
  if (yb < 0) yb = 0;
  if (ye > 1.0) ye = 1.0;
  xminx0 = xminx = xb * g.xsize;
  xmaxx = xe * g.xsize;
  xresx = (xmaxx - xminx) / (double)(xresol - 1);
  yminy0 = yminy = yb * g.ysize;
  ymaxy = ye * g.ysize;
  yresy = (ymaxy - yminy) / (double)(yresol - 1);
  yminy -= yresy;
  if (yminy < 0) yminy = 0;
  ymaxy += yresy;
  if (ymaxy > g.ysize) ymaxy = g.ysize;
  /* one chunk: xresolsum = 0 (loadjmart.c:701-708) */
  xminx -= xresx;
  if (xminx < 0) xminx = 0;
  xmaxx = xb * g.xsize + xresx * (double)xresol;
  if (xmaxx > g.xsize) xmaxx = g.xsize;
  ival = 0.1 * (double)M / (double)nresol;
  markstp = (int64_t)ival;
  if (markstp < 1) markstp = 1;
  val0 = (float *)calloc(2 * nresol, sizeof(float));
  lin0 = (int64_t *)calloc(2 * nresol, sizeof(int64_t));
  esp = (double *)calloc(X * Y, sizeof(double));
  if (!val0 || !lin0 || !esp) return I2E_ERR_ARG;
  for (m1 = 1; m1 < X - 1; m1++)
    for (m2 = 1; m2 < Y - 1; m2++) esp[m1 * Y + m2] = spncalc(&g, s, m1, m2);
  if (melt) memset(melt, 0, sizeof(double) * M);
#define INWIN(k) ((double)s->markx[k] > xminx && (double)s->markx[k] < xmaxx && (double)s->marky[k] > yminy && \
                  (double)s->marky[k] < ymaxy)
#define POS(k)                                                    \
  do {                                                            \
    ea = ((double)(s->markx[k]) - xminx0) / xresx + 1.0;          \
    m1 = (int64_t)(ea);                                           \
    if (m1 < 0) m1 = 0;                                           \
    if (m1 > xresol) m1 = xresol;                                 \
    ea -= (double)(m1);                                           \
    na = ((double)(s->marky[k]) - yminy0) / yresy + 1.0;          \
    m2 = (int64_t)(na);                                           \
    if (m2 < 0) m2 = 0;                                           \
    if (m2 > yresol) m2 = yresol;                                 \
    na -= (double)(m2);                                           \
  } while (0)
  if (type == 5) {
    int cmin = 1000, cmax = 0, ccur, yn;
    for (m3 = 0; m3 < M; m3 += markstp)
      if (INWIN(m3) && s->markt[m3] < 50) {
        if (cmin > s->markt[m3]) cmin = s->markt[m3];
        if (cmax < s->markt[m3]) cmax = s->markt[m3];
      }
    for (ccur = cmin; ccur <= cmax; ccur++) {
      for (m3 = 0; m3 < M; m3 += markstp)
        if (INWIN(m3) && s->markt[m3] == ccur) {
          POS(m3);
          m4 = nresol + (m1 - 1) * yresol + (m2 - 1);
          if (m1 > 0 && m2 > 0) val0[m4] += (float)((1.0 - ea) * (1.0 - na));
          if (m1 > 0 && m2 < yresol) val0[m4 + 1] += (float)((1.0 - ea) * na);
          if (m1 < xresol && m2 > 0) val0[m4 + yresol] += (float)(ea * (1.0 - na));
          if (m1 < xresol && m2 < yresol) val0[m4 + yresol + 1] += (float)(ea * na);
        }
      for (m1 = 0; m1 < nresol; m1++)
        if (val0[nresol + m1] > val0[m1]) {
          val0[m1] = val0[nresol + m1];
          lin0[m1] = ccur;
          val0[nresol + m1] = 0;
        }
    }
    yn = 0;
    for (m1 = 0; m1 < nresol; m1++)
      if (!val0[m1]) {
        lin0[m1] = -1;
        yn = 1;
      }
    if (sp->gridmod && yn) {
      int64_t m5, m6, m7, r;
      for (m1 = 0; m1 < xresol; m1++)
        for (m2 = 0; m2 < yresol; m2++) {
          m3 = m1 * yresol + m2;
          if (lin0[m3] != -1) continue;
          r = 1;
          yn = 0;
          do {
            cmin = 1000, cmax = 0;
            for (m5 = m1 - r; m5 <= m1 + r; m5++)
              for (m6 = m2 - r; m6 <= m2 + r; m6++)
                if (m5 >= 0 && m5 < xresol && m6 >= 0 && m6 < yresol) {
                  m7 = m5 * yresol + m6;
                  if (lin0[m7] != -1) {
                    yn = 1;
                    if (cmin > lin0[m7]) cmin = (int)lin0[m7];
                    if (cmax < lin0[m7]) cmax = (int)lin0[m7];
                  }
                }
            r++;
          } while (!yn && r <= sp->gridmod);
          if (!yn) continue;
          for (ccur = cmin; ccur <= cmax; ccur++) {
            for (m5 = m1 - r + 1; m5 < m1 + r; m5++)
              for (m6 = m2 - r + 1; m6 < m2 + r; m6++)
                if (m5 >= 0 && m5 < xresol && m6 >= 0 && m6 < yresol) {
                  m7 = m5 * yresol + m6;
                  if (lin0[m7] == ccur) {
                    ea = ABSV(((double)(m5 - m1)) / (double)(r));
                    na = ABSV(((double)(m6 - m2)) / (double)(r));
                    val0[nresol + m3] += (float)((1.0 - ea) * (1.0 - na)) * val0[m7];
                  }
                }
            if (val0[nresol + m3] > val0[m3]) {
              val0[m3] = val0[nresol + m3];
              lin0[nresol + m3] = ccur;
              val0[nresol + m3] = 0;
            }
          }
        }
      for (m1 = 0; m1 < nresol; m1++)
        if (lin0[m1] == -1 && val0[m1]) {
          lin0[m1] = lin0[nresol + m1];
          lin0[nresol + m1] = 0;
        }
    }
    for (m1 = 0; m1 < nresol; m1++) rock[m1] = (int32_t)lin0[m1];
  } else {
    for (m3 = 0; m3 < M; m3 += markstp)
      if (INWIN(m3) && s->markt[m3] < 50) {
        E1 e;
        double mpb, mtk, mnu, mro, mbb, mcp, mkt, mht;
        int mm2;
        allinter1(&g, s, esp, (double)s->markx[m3], (double)s->marky[m3], &e);
        mpb = e.pr * 1e-5;
        if (s->markk[m3] < 1.0) s->markk[m3] = (float)(e.t);
        mtk = (double)(s->markk[m3]);
        if (s->markt[m3] != 0)
          if (s->marktmax[m3] < s->markk[m3]) s->marktmax[m3] = (float)(s->markk[m3]);
        mm2 = s->markt[m3];
        if (melt) {
          double xm = 0;
          const int base = mm2 < 20 ? mm2 : mm2 - 20;
          if (p->melt_on && base >= 0 && base < I2E_MAXTYPES && p->meltable[base]) {
            xm = (mtk - p->melt_ts) / (p->melt_tl - p->melt_ts);
            if (xm < 0.0) xm = 0.0;
            if (xm > 1.0) xm = 1.0;
          }
          melt[m3] = xm;
        }
        mnu = s->markv[m3];
        if (mnu < p->nubeg) mnu = p->nubeg;
        if (mnu > p->nuend) mnu = p->nuend;
        mro = dencalc(p, mtk, mpb, (double)(s->markpor[m3]), mm2, &mbb);
        mcp = p->rock[mm2].cp;
        mkt = (p->rock[mm2].kt + p->rock[mm2].kf / (mtk + 77.0)) * exp(p->rock[mm2].kp * mpb);
        if (p->rock[mm2].kt < 0)
          mkt = -p->rock[mm2].kt / (1.0 + p->rock[mm2].kf * (mtk - p->rock[mm2].kp) / p->rock[mm2].kp);
        mht = 0.0; /* markht: not carried */
        (void)mcp;
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
        POS(m3);
        m4 = (m1 - 1) * yresol + (m2 - 1);
        if (m1 > 0 && m2 > 0) {
          val0[m4] += (float)(ival * (1.0 - ea) * (1.0 - na));
          val0[nresol + m4] += (float)((1.0 - ea) * (1.0 - na));
        }
        if (m1 > 0 && m2 < yresol) {
          val0[m4 + 1] += (float)(ival * (1.0 - ea) * na);
          val0[nresol + m4 + 1] += (float)((1.0 - ea) * na);
        }
        if (m1 < xresol && m2 > 0) {
          val0[m4 + yresol] += (float)(ival * ea * (1.0 - na));
          val0[nresol + m4 + yresol] += (float)(ea * (1.0 - na));
        }
        if (m1 < xresol && m2 < yresol) {
          val0[m4 + yresol + 1] += (float)(ival * ea * na);
          val0[nresol + m4 + yresol + 1] += (float)(ea * na);
        }
      }
    for (m1 = 0; m1 < xresol; m1++)
      for (m2 = 0; m2 < yresol; m2++) {
        m4 = m1 * yresol + m2;
        if (val0[nresol + m4]) {
          val0[m4] /= val0[nresol + m4];
        } else {
          E1 e;
          allinter1(&g, s, esp, xminx0 + xresx * ((double)(m1)), yminy0 + yresy * ((double)(m2)), &e);
          switch (type) {
            case 1: val0[m4] = (float)(e.t); break;
            case 2: val0[m4] = (float)(log(e.nu) / log(10.0)); break;
            case 4:
              val0[m4] = (float)(2.0 * e.exysxy + e.exxsxx + e.eyysyy);
              if (val0[m4] < 1e-30) val0[m4] = 1e-30;
              val0[m4] = (float)(log(val0[m4]) / log(10.0));
              break;
            case 6: val0[m4] = (float)(e.ro); break;
            case 7: val0[m4] = (float)(e.kt); break;
            case 8: val0[m4] = (float)(e.ht + 2.0 * e.exysxy + e.exxsxx + e.eyysyy); break;
            case 12:
              ival = pow(e.exy * e.exy + 0.5 * (e.exx * e.exx + e.eyy * e.eyy), 0.5);
              if (ival < 1e-30) ival = 1e-30;
              val0[m4] = e.esp / ival;
              break;
          }
        }
      }
    memcpy(val, val0, sizeof(float) * nresol);
  }
#undef INWIN
#undef POS
  free(val0);
  free(lin0);
  free(esp);
  return I2E_OK;
}

/* grain(), grain.c:4-156: olivine grain growth and the pallasite emplacement
 * after the first impact, every marker in index order; markgr updated in
 * place, types of s changed                                                 */
int orc_grain(const i2e_params *p, i2e_host_state *s, const i2e_grain_params *gp, double timesum, double timestep,
              float *markgr) {
  double n_grain = 3.70, E_grain = 1.0098e+5, k0_grain = 3.2033, n_oliv = 2.250, E_oliv = 5.71674e+5,
         k0_oliv = 1.06593e+16, pival = 3.141592654, radius_core = 1.000e+5, radius_planet = 2.000e+5,
         ref_angle = 45.000, dev_angle = 10.000, time_window = 0.45;
  double curr_angle, distancee, dist_x, dist_y, a, b, c;
  const int growth_model = gp->growth_model;
  int64_t mm1;
  for (mm1 = 0; mm1 < s->marknum; mm1++) {
    if ((growth_model == 1 || growth_model == 2) &&
        ((timesum + timestep) >= (gp->impact_time_ma * 1.000e+6 * 365.250 * 24.000 * 3600.000)) &&
        ((timesum + timestep) <= ((gp->impact_time_ma + time_window) * 1.000e+6 * 365.250 * 24.000 * 3600.000))) {
      dist_x = s->markx[mm1] - (p->xsize / 2.000);
      dist_y = s->marky[mm1] - (p->ysize / 2.000);
      distancee = pow((dist_x * dist_x + dist_y * dist_y), 0.500);
      if (distancee >= radius_core && distancee <= radius_planet && dist_x > 0.000 && dist_y < 0.000) {
        curr_angle = asin(dist_x / distancee) / pival * 180.000;
        if (curr_angle >= ref_angle - dev_angle && curr_angle <= ref_angle + dev_angle) {
          if (s->markt[mm1] == 6) s->markt[mm1] = 5;
          if (s->markt[mm1] == 26) s->markt[mm1] = 25;
          markgr[mm1] = (float)(gp->gr_init);
        }
      }
    }
    if ((timesum + timestep) >= (gp->impact_time_ma * 1.000e+6 * 365.250 * 24.000 * 3600.000)) {
      if ((s->markt[mm1] == 5 && s->markk[mm1] >= 1073.000) || (s->markt[mm1] == 25 && s->markk[mm1] < 2103.000)) {
        a = pow(((double)(markgr[mm1])), n_grain);
        b = k0_grain * exp(-E_grain / (8.3145 * s->markk[mm1])) * timestep;
        c = pow((a + b), (1.000 / n_grain));
        markgr[mm1] = (float)(c);
      }
    }
    if ((growth_model == 1 && (timesum + timestep) >= (gp->impact_time_ma * 1.000e+6 * 365.250 * 24.000 * 3600.000)) ||
        (growth_model == 2 && (timesum + timestep) >= (gp->start_time_ma * 1.000e+6 * 365.250 * 24.000 * 3600.000))) {
      if (s->markt[mm1] == 6 || (s->markt[mm1] == 26 && s->markk[mm1] < 2103.000)) {
        a = pow(((double)(markgr[mm1])), n_oliv);
        b = k0_oliv * exp(-E_oliv / (8.3145 * s->markk[mm1])) * timestep;
        c = pow((a + b), (1.000 / n_oliv));
        markgr[mm1] = (float)(c);
      }
      if ((s->markt[mm1] == 25 || s->markt[mm1] == 26) && s->markk[mm1] >= 2103.000) markgr[mm1] = (float)(gp->gr_init);
    }
  }
  return I2E_OK;
}
