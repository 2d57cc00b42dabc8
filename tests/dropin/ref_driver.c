/* ref_driver.c -- a stand-in for the reference's main() (i2mart.c:60-164)
 * that drives the drop-in step functions of include/i2e_dropin.h over
 * globals named and typed as headmart.c defines them.
 *
 *   ref_driver <state.prn> <mode.txt> <nsteps> <out.bin>
 *
 * The state comes from a binary .prn file (the reference's loader format,
 * read with i2e_prn_read straight into the globals); mode.txt holds
 * "name value" lines for the run parameters the reference reads from
 * mode.t3c, plus the build's hook values (radmax_frac, longlived,
 * porosity_on, solver tolerances).  After one ronurecalc() the loop body of
 * i2mart.c:60-164 runs nsteps times (movemod, tempmod, markmod, gridmod on,
 * timedir 1, no ratemod, no impacts).  out.bin receives the node fields vx,
 * vy, pr, tk, gp (FP64), the markers x, y, T (float32), types (int8), then
 * marknum, timestep, timesum and GXKOEF.                                     */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "i2e_dropin.h"

#define DMAXNOD 70000
#define DMAXMRK 1500000
#define DMAXBON (4 * DMAXNOD)
#define DMAXTMR 200

/* the reference's globals (headmart.c), sized for the test cases */
double gx[1201], gy[1201];
double vx[DMAXNOD], vy[DMAXNOD], pr[DMAXNOD], gp[DMAXNOD], ga[DMAXNOD];
double exx[DMAXNOD], eyy[DMAXNOD], exy[DMAXNOD], sxx[DMAXNOD], syy[DMAXNOD], sxy[DMAXNOD];
double nu[DMAXNOD], nd[DMAXNOD], ro[DMAXNOD], tk[DMAXNOD], cp[DMAXNOD], kt[DMAXNOD], ht[DMAXNOD], et[DMAXNOD];
float markx[DMAXMRK], marky[DMAXMRK], markk[DMAXMRK], markv[DMAXMRK], markp[DMAXMRK], markpor[DMAXMRK],
    marktmax[DMAXMRK];
char markt[DMAXMRK], markmg_old[DMAXMRK];
double markmg_time[DMAXMRK];
long int bondm[DMAXBON], bondn[DMAXBON][3];
double bondv[DMAXBON][4];
long int xnumx, ynumy, mnumx, mnumy, marknum, bondnum;
double marknu[DMAXTMR], markdh[DMAXTMR], markdv[DMAXTMR], markss[DMAXTMR], markmm[DMAXTMR], markn0[DMAXTMR],
    markn1[DMAXTMR];
double markro[DMAXTMR], markaa[DMAXTMR], markbb[DMAXTMR], markcp[DMAXTMR], markkt[DMAXTMR], markkf[DMAXTMR],
    markkp[DMAXTMR];
int markim[DMAXTMR], rocknum;
double xsize, ysize, GXKOEF, GYKOEF, tmp_ambient, al2627_init, fe6056_init;
double maxxystep, maxtkstep, maxtmstep, timestep, timesum;
double nubeg, nuend, nucontr, strmin, heatdif, frictyn, adiabyn, por_init, k_cutoff;
int heatmod, densimod, outgrid, markmod;
double core_form_time;
int dynamo;

static double radmax_frac = 0.95, longlived = 1.0, h0_scale = 1.0, stokes_rtol = 1e-10, heat_rtol = 1e-12,
              grav_rtol = 1e-12;
static int porosity_on = 1;
static double fe_ts = 0, fe_tl = 0;

static void hook(i2e_params *p, i2e_solver_opts *o) {
  p->radmax_frac = radmax_frac;
  p->longlived = longlived;
  p->h0_scale = h0_scale;
  p->porosity_on = porosity_on;
  p->fe_ts = fe_ts;
  p->fe_tl = fe_tl;
  o->stokes_rtol = stokes_rtol;
  o->heat_rtol = heat_rtol;
  o->grav_rtol = grav_rtol;
}

static void read_mode(const char *path) {
  char name[64];
  double v;
  FILE *f = fopen(path, "r");
  if (!f) {
    perror(path);
    exit(20);
  }
  while (fscanf(f, "%63s %lf", name, &v) == 2) {
#define D(x) else if (!strcmp(name, #x)) x = v;
#define I(x) else if (!strcmp(name, #x)) x = (int)v;
    if (0) {
    }
    D(maxxystep) D(maxtkstep) D(maxtmstep) D(nubeg) D(nuend) D(nucontr) D(strmin) D(heatdif) D(frictyn)
    D(adiabyn) D(por_init) D(k_cutoff) D(timestep) D(radmax_frac) D(longlived) D(h0_scale) D(stokes_rtol)
    D(heat_rtol) D(grav_rtol) I(heatmod) I(densimod) I(outgrid) I(markmod) I(porosity_on) D(core_form_time)
    I(dynamo) D(fe_ts) D(fe_tl)
    else {
      fprintf(stderr, "unknown mode parameter %s\n", name);
      exit(21);
    }
#undef D
#undef I
  }
  fclose(f);
}

static void load(const char *path) {
  i2e_prn_header h;
  i2e_params p;
  i2e_host_state s;
  int n0, t;
  if (i2e_prn_read_header(path, &h)) {
    fprintf(stderr, "%s\n", i2e_last_error());
    exit(22);
  }
  if (h.xnumx * h.ynumy > DMAXNOD || h.marknum > DMAXMRK || h.bondnum > DMAXBON || h.xnumx > 1201 ||
      h.ynumy > 1201) {
    fprintf(stderr, "state too large for the driver\n");
    exit(23);
  }
  memset(&p, 0, sizeof p);
  memset(&s, 0, sizeof s);
  s.gx = gx, s.gy = gy, s.vx = vx, s.vy = vy, s.pr = pr, s.gp = gp, s.ga = ga;
  s.exx = exx, s.eyy = eyy, s.exy = exy, s.sxx = sxx, s.syy = syy, s.sxy = sxy;
  s.nu = nu, s.nd = nd, s.ro = ro, s.tk = tk, s.cp = cp, s.kt = kt, s.ht = ht, s.et = et;
  s.marknum = h.marknum;
  s.markx = markx, s.marky = marky, s.markk = markk, s.markv = markv, s.markp = markp;
  s.markpor = markpor, s.marktmax = marktmax, s.markt = (int8_t *)markt;
  s.markmg_old = (int8_t *)markmg_old, s.markmg_time = markmg_time;
  s.bondm = (int64_t *)bondm, s.bondn = (int64_t *)&bondn[0][0], s.bondv = &bondv[0][0];
  s.bondnum = h.bondnum;
  if (i2e_prn_read(path, &p, &s, &n0)) {
    fprintf(stderr, "%s\n", i2e_last_error());
    exit(24);
  }
  xnumx = p.xnumx, ynumy = p.ynumy, mnumx = p.mnumx, mnumy = p.mnumy;
  marknum = s.marknum, bondnum = s.bondnum;
  xsize = p.xsize, ysize = p.ysize, GYKOEF = p.gykoef, GXKOEF = s.gxkoef;
  tmp_ambient = p.tmp_ambient, al2627_init = p.al2627_init, fe6056_init = p.fe6056_init;
  timesum = s.timesum;
  rocknum = p.ntypes;
  for (t = 0; t < rocknum; t++) {
    const i2e_rock *r = &p.rock[t];
    marknu[t] = r->nu0, markdh[t] = r->dh, markdv[t] = r->dv, markss[t] = r->ss, markmm[t] = r->mm;
    markn0[t] = r->n0, markn1[t] = r->n1, markro[t] = r->ro, markaa[t] = r->aa, markbb[t] = r->bb;
    markcp[t] = r->cp, markkt[t] = r->kt, markkf[t] = r->kf, markkp[t] = r->kp, markim[t] = 0;
  }
}

int main(int argc, char **argv) {
  int n0, nsteps;
  long int N;
  FILE *f;
  if (argc != 5) {
    fprintf(stderr, "usage: %s state.prn mode.txt nsteps out.bin\n", argv[0]);
    return 2;
  }
  load(argv[1]);
  read_mode(argv[2]);
  nsteps = atoi(argv[3]);
  i2e_dropin_set_hook(hook);
  ronurecalc();
  for (n0 = 0; n0 < nsteps; n0++) { /* i2mart.c:60-153 */
    timestep = maxtmstep;
    maxvelstep();
    viterate(n0);
    timestep = maxtmstep;
    maxvelstep();
    if (timestep) titerate(n0);
    if (markmod && timestep) movemark();
    ronurecalc();
    timesum += timestep;
    fprintf(stderr, "step %d dt %.6e\n", n0, timestep);
  }
  i2e_dropin_release();
  N = xnumx * ynumy;
  f = fopen(argv[4], "wb");
  if (!f) {
    perror(argv[4]);
    return 25;
  }
  fwrite(vx, sizeof(double), N, f);
  fwrite(vy, sizeof(double), N, f);
  fwrite(pr, sizeof(double), N, f);
  fwrite(tk, sizeof(double), N, f);
  fwrite(gp, sizeof(double), N, f);
  fwrite(markx, sizeof(float), marknum, f);
  fwrite(marky, sizeof(float), marknum, f);
  fwrite(markk, sizeof(float), marknum, f);
  fwrite(markt, 1, marknum, f);
  fwrite(markmg_old, 1, marknum, f);
  fwrite(markmg_time, sizeof(double), marknum, f);
  fwrite(&marknum, sizeof marknum, 1, f);
  fwrite(&timestep, sizeof timestep, 1, f);
  fwrite(&timesum, sizeof timesum, 1, f);
  fwrite(&GXKOEF, sizeof GXKOEF, 1, f);
  fclose(f);
  return 0;
}
