/*
 * i2e_dropin.h -- the reference's own step functions, same names and
 * signatures, over the reference's global arrays (headmart.c:87-132,
 * 170-259), running on the GPU through libi2e.
 *
 * A maintainer compiles i2elvis_planet_amd/csrc/dropin.c into the reference
 * program in place of the definitions it replaces and links libi2e.so:
 *
 *   void maxvelstep(void)   replaces movemart.c:1225-1262
 *   void viterate(int n0)   replaces movemart.c:603-1217 (+ its gausmat4
 *                           PARDISO solves, gausmart.c:174-318)
 *   void titerate(int n0)   replaces heatmart.c:320-695
 *   void movemark(void)     replaces markmart.c:3-188
 *   void ronurecalc(void)   replaces markmart.c:192-1019 (+ tkrecalc,
 *                           heatmart.c:758-826)
 *
 * main() (i2mart.c:60-164) keeps calling them in its own order; impact()
 * and pebbleaccr() still run between movemark() and ronurecalc() on the
 * host globals.  Each call copies the globals the step reads to the GPU,
 * runs the step there and copies back everything it writes (node fields,
 * markers, marknum, timestep, timesum, GXKOEF), so host code between the
 * calls sees and may change the state exactly as with the reference's
 * functions.  Per call that is about 29 B per marker and 152 B per node
 * each way over PCIe.
 *
 * The globals bound (names and types as headmart.c defines them):
 *   double gx[], gy[], vx[], vy[], pr[], gp[], ga[], exx[], eyy[], exy[],
 *          sxx[], syy[], sxy[], nu[], nd[], ro[], tk[], cp[], kt[], ht[], et[];
 *   float  markx[], marky[], markk[], markv[], markp[], markpor[], marktmax[];
 *   char   markt[], markmg_old[];   double markmg_time[];
 *   long   bondm[], bondn[][3];  double bondv[][4];
 *   long   xnumx, ynumy, mnumx, mnumy, marknum, bondnum;
 *   double marknu[], markdh[], markdv[], markss[], markmm[], markn0[], markn1[],
 *          markro[], markaa[], markbb[], markcp[], markkt[], markkf[], markkp[];
 *   int    markim[], rocknum;
 *   double xsize, ysize, GXKOEF, GYKOEF, tmp_ambient, al2627_init, fe6056_init,
 *          maxxystep, maxtkstep, maxtmstep, timestep, timesum, nubeg, nuend,
 *          nucontr, strmin, heatdif, frictyn, adiabyn, por_init, k_cutoff;
 *   int    heatmod, densimod, outgrid, markmod, dynamo;   double core_form_time;
 *
 * Errors: like the reference (gausmart.c:271-304, exit on a failed solve),
 * a failing call prints the library's message to stderr and exits the
 * process, with the i2e status code (include/i2e.h) as exit status.
 */
#ifndef I2E_DROPIN_H
#define I2E_DROPIN_H

#include "i2e.h"

#ifdef __cplusplus
extern "C" {
#endif

void maxvelstep(void);
void viterate(int n0);
void titerate(int n0);
void movemark(void);
void ronurecalc(void);

/* Optional (no counterpart in the reference): a hook called before every
 * step call with the parameters and solver options the shim built from the
 * globals, so a caller can change what the reference fixes in code
 * (radmax 0.95 xsize/2 and 1e22 Pa s, markmart.c:268, 340; the long-lived
 * isotopes, markmart.c:772-782; the porosity closures of types 5 and 6) or
 * the solver (default: FP64, Stokes 1e-10, heat and gravity 1e-12, close to
 * the reference's direct solves).  NULL removes the hook. */
typedef void (*i2e_dropin_hook)(i2e_params *p, i2e_solver_opts *o);
void i2e_dropin_set_hook(i2e_dropin_hook hook);
/* Release the GPU context (the next call creates a new one). */
void i2e_dropin_release(void);

#ifdef __cplusplus
}
#endif
#endif /* I2E_DROPIN_H */
