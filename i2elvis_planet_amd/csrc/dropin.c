/* dropin.c -- the reference's step functions by name (include/i2e_dropin.h)
 * over the reference's globals (headmart.c:87-259), each call running the
 * step on the GPU through libi2e's context API.
 *
 * Compiled into the reference program (it refers to the globals headmart.c
 * defines) and linked with libi2e.so; no GPU code lives here. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "i2e_dropin.h"

/* ---- the reference's globals (headmart.c) ------------------------------ */
extern double gx[], gy[], vx[], vy[], pr[], gp[], ga[], exx[], eyy[], exy[], sxx[], syy[], sxy[];
extern double nu[], nd[], ro[], tk[], cp[], kt[], ht[], et[];
extern float markx[], marky[], markk[], markv[], markp[], markpor[], marktmax[];
extern char markt[], markmg_old[];
extern double markmg_time[];
extern long int bondm[], bondn[][3];
extern double bondv[][4];
extern long int xnumx, ynumy, mnumx, mnumy, marknum, bondnum;
extern double marknu[], markdh[], markdv[], markss[], markmm[], markn0[], markn1[];
extern double markro[], markaa[], markbb[], markcp[], markkt[], markkf[], markkp[];
extern int markim[], rocknum;
extern double xsize, ysize, GXKOEF, GYKOEF, tmp_ambient, al2627_init, fe6056_init;
extern double maxxystep, maxtkstep, maxtmstep, timestep, timesum;
extern double nubeg, nuend, nucontr, strmin, heatdif, frictyn, adiabyn, por_init, k_cutoff;
extern int heatmod, densimod, outgrid, markmod;
extern double core_form_time;
extern int dynamo;

static i2e_ctx *g_ctx = NULL;
static int64_t g_cap = 0, g_X = 0, g_Y = 0, g_mx = 0, g_my = 0;
static i2e_dropin_hook g_hook = NULL;

void i2e_dropin_set_hook(i2e_dropin_hook hook) { g_hook = hook; }

void i2e_dropin_release(void) {
  if (g_ctx) i2e_destroy(g_ctx);
  g_ctx = NULL;
  g_cap = 0;
}

static void die(const char *fn, int st) {
  fprintf(stderr, "%s: %s (i2e status %d)\n", fn, i2e_last_error(), st);
  exit(st);
}

/* i2e_params from the globals, with the constants the reference fixes in
 * code: radmax = 0.95 xsize/2 with 1e22 Pa s beyond (markmart.c:268, 340),
 * the long-lived isotopes (markmart.c:772-782), porosity closures for the
 * solid silicate types 5 and 6 (markmart.c:347-360, 612-648), 26Al in types
 * 5, 6, 25, 26 and 60Fe in 7-10, 17-19 (markmart.c:760, 789), the
 * soft-turbulence conductivity after core_form_time (markmart.c:508-610)
 * and the magnetisation of solid silicates (markmart.c:651-744).  The
 * reference's iron solidus and liquidus depend on pressure (markmart.c:
 * 2519-2560); the shim leaves fe_tl = 0, iron solid, unless the hook sets
 * them.                                                                   */
static void build_params(i2e_params *p) {
  int t;
  memset(p, 0, sizeof *p);
  p->xnumx = xnumx;
  p->ynumy = ynumy;
  p->mnumx = mnumx;
  p->mnumy = mnumy;
  p->xsize = xsize;
  p->ysize = ysize;
  p->maxxystep = maxxystep;
  p->maxtkstep = maxtkstep;
  p->maxtmstep = maxtmstep;
  p->nubeg = nubeg;
  p->nuend = nuend;
  p->nucontr = nucontr;
  p->strmin = strmin;
  p->heatdif = heatdif;
  p->frictyn = frictyn;
  p->adiabyn = adiabyn;
  p->heatmod = heatmod;
  p->densimod = densimod;
  p->outgrid = outgrid;
  p->markmod = markmod == 1 ? 1 : 2; /* markmart.c:74: 1 = Euler, else RK4 */
  p->gykoef = GYKOEF;
  p->tmp_ambient = tmp_ambient;
  p->al2627_init = al2627_init;
  p->fe6056_init = fe6056_init;
  p->h0_scale = 1.0;
  p->longlived = 1.0;
  p->por_init = por_init;
  p->k_cutoff = k_cutoff;
  p->radmax_frac = 0.95;
  p->eta_outer = 1.0e22;
  p->melt_on = 0;
  p->porosity_on = 1;
  p->soft_turb = 1;
  p->core_form_time = core_form_time;
  p->magnet = 1;
  p->dynamo = dynamo;
  p->ntypes = rocknum < I2E_MAXTYPES ? rocknum : I2E_MAXTYPES;
  for (t = 0; t < p->ntypes; t++) {
    i2e_rock *r = &p->rock[t];
    r->nu0 = marknu[t];
    r->dh = markdh[t];
    r->dv = markdv[t];
    r->ss = markss[t];
    r->mm = markmm[t];
    r->n0 = markn0[t];
    r->n1 = markn1[t];
    r->ro = markro[t];
    r->aa = markaa[t];
    r->bb = markbb[t];
    r->cp = markcp[t];
    r->kt = markkt[t];
    r->kf = markkf[t];
    r->kp = markkp[t];
    p->immobile[t] = markim[t] != 0;
  }
  p->is_silicate[5] = p->is_silicate[6] = p->is_silicate[25] = p->is_silicate[26] = 1;
  for (t = 7; t <= 10; t++) p->is_iron[t] = 1;
  for (t = 17; t <= 19; t++) p->is_iron[t] = 1;
  p->porous[5] = p->porous[6] = 1;
}

static void bind_state(i2e_host_state *s) {
  memset(s, 0, sizeof *s);
  s->gx = gx;
  s->gy = gy;
  s->vx = vx;
  s->vy = vy;
  s->pr = pr;
  s->gp = gp;
  s->ga = ga;
  s->exx = exx;
  s->eyy = eyy;
  s->exy = exy;
  s->sxx = sxx;
  s->syy = syy;
  s->sxy = sxy;
  s->nu = nu;
  s->nd = nd;
  s->ro = ro;
  s->tk = tk;
  s->cp = cp;
  s->kt = kt;
  s->ht = ht;
  s->et = et;
  s->marknum = marknum;
  s->markx = markx;
  s->marky = marky;
  s->markk = markk;
  s->markv = markv;
  s->markp = markp;
  s->markpor = markpor;
  s->marktmax = marktmax;
  s->markt = (int8_t *)markt;
  s->markmg_old = (int8_t *)markmg_old;
  s->markmg_time = markmg_time;
  s->bondm = (int64_t *)bondm;
  s->bondn = (int64_t *)&bondn[0][0];
  s->bondv = &bondv[0][0];
  s->bondnum = bondnum;
  s->timestep = timestep;
  s->timesum = timesum;
  s->gxkoef = GXKOEF;
}

/* Globals -> GPU before a step call (the context is made, or remade when the
 * grid changed or the markers outgrew it). */
static void enter(const char *fn) {
  i2e_params p;
  i2e_solver_opts o;
  i2e_host_state s;
  int st;
  if (sizeof(long int) != sizeof(int64_t)) die(fn, I2E_ERR_UNSUPPORTED);
  if (rocknum > I2E_MAXTYPES) {
    fprintf(stderr, "%s: %d rock types, at most %d supported\n", fn, rocknum, I2E_MAXTYPES);
    exit(I2E_ERR_CAPACITY);
  }
  build_params(&p);
  i2e_default_solver(&o);
  o.precision = I2E_PREC_FP64;
  o.stokes_rtol = 1e-10;
  o.heat_rtol = 1e-12;
  o.grav_rtol = 1e-12;
  if (g_hook) g_hook(&p, &o);
  if (g_ctx && (g_X != xnumx || g_Y != ynumy || g_mx != mnumx || g_my != mnumy || marknum > g_cap))
    i2e_dropin_release();
  if (!g_ctx) {
    int64_t cap = marknum + marknum / 4 + 1024;
    if ((st = i2e_create(&p, cap, &g_ctx))) die(fn, st);
    g_cap = cap;
    g_X = xnumx;
    g_Y = ynumy;
    g_mx = mnumx;
    g_my = mnumy;
  } else if ((st = i2e_set_params(g_ctx, &p))) {
    die(fn, st);
  }
  if ((st = i2e_set_solver(g_ctx, &o))) die(fn, st);
  bind_state(&s);
  if ((st = i2e_upload(g_ctx, &s))) die(fn, st);
}

/* GPU -> globals after a step call. */
static void leave(const char *fn) {
  i2e_host_state s;
  int st;
  bind_state(&s);
  if ((st = i2e_download(g_ctx, &s))) die(fn, st);
  marknum = s.marknum;
  timestep = s.timestep;
  timesum = s.timesum;
  GXKOEF = s.gxkoef;
}

void maxvelstep(void) {
  int st;
  enter("maxvelstep");
  if ((st = i2e_maxvelstep(g_ctx, NULL))) die("maxvelstep", st);
  leave("maxvelstep");
}

void viterate(int n0) {
  int st;
  enter("viterate");
  if ((st = i2e_viterate(g_ctx, n0))) die("viterate", st);
  leave("viterate");
}

void titerate(int n0) {
  int st;
  enter("titerate");
  if ((st = i2e_titerate(g_ctx, n0))) die("titerate", st);
  leave("titerate");
}

void movemark(void) {
  int st;
  enter("movemark");
  if ((st = i2e_movemark(g_ctx))) die("movemark", st);
  leave("movemark");
}

void ronurecalc(void) {
  int st;
  enter("ronurecalc");
  if ((st = i2e_ronurecalc(g_ctx))) die("ronurecalc", st);
  leave("ronurecalc");
}
