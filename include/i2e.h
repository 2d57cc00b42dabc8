/*
 * i2e.h -- C-ABI of the MI355X-native I2ELVIS marker-in-cell time step.
 *
 * The five step functions replace the reference's serial C entry points
 * (prototypes in headmart.c:321-366, called in fixed order by main() in
 * i2mart.c:60-164):
 *
 *   i2e_maxvelstep  <- void maxvelstep(void)   movemart.c:1225-1262
 *   i2e_viterate    <- void viterate(int)      movemart.c:603-1217
 *                      (+ the three gausmat4(0,..) PARDISO solves it makes,
 *                       gausmart.c:174-318, replaced by matrix-free solvers)
 *   i2e_titerate    <- void titerate(int)      heatmart.c:320-695
 *   i2e_movemark    <- void movemark(void)     markmart.c:3-188
 *   i2e_ronurecalc  <- void ronurecalc(void)   markmart.c:192-1019
 *                      (+ tkrecalc/gausmat3,   heatmart.c:758-826)
 *
 * The reference keeps all state in global arrays (headmart.c:38-167); here the
 * caller owns host arrays described by i2e_host_state, and a context owns the
 * device copies.  Index maps are the reference's: node m = m1*ynumy + m2 (y
 * fastest); Stokes unknown 3m+{0:P,1:Vx,2:Vy}; temperature boundary slot
 * 3N+m (movemart.c:800, heatmart.c:355).  Markers are float32 SoA plus a
 * signed char type, nodes are FP64 (headmart.c:87-127).
 *
 * Every function returns an int status (I2E_OK == 0).  Nothing here exits the
 * process; a caller that wants the reference's exit() behaviour maps the codes.
 */
#ifndef I2E_H
#define I2E_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  I2E_OK = 0,
  I2E_ERR_ARG = 1,          /* bad argument / inconsistent sizes            */
  I2E_ERR_CAPACITY = 2,     /* size above a compiled-in maximum             */
  I2E_ERR_HIP = 3,          /* a HIP runtime call failed                    */
  I2E_ERR_NOCONV = 4,       /* a Krylov solve missed its tolerance          */
  I2E_ERR_UNSUPPORTED = 5,  /* input valid for the reference, not handled   */
  I2E_ERR_NODEVICE = 6      /* no GPU visible                               */
};

#define I2E_MAXTYPES 64     /* rock types (reference MAXTMR=200, 0..63 used) */

/* ---- per-rock-type constants (reference: marknu..markkp, headmart.c:127-131,
 *      read from the ROCKS table of init.t3c) --------------------------------*/
typedef struct i2e_rock {
  double nu0;   /* marknu: viscosity prefactor, Pa s                        */
  double dh;    /* markdh: activation energy, J/mol                         */
  double dv;    /* markdv: activation volume, J/bar                         */
  double ss;    /* markss: transition stress (0 = Newtonian)                */
  double mm;    /* markmm: stress exponent                                  */
  double n0;    /* markn0: lower viscosity limit; molten types return this  */
  double n1;    /* markn1: upper viscosity limit                            */
  double ro;    /* markro: reference density, kg/m^3                        */
  double aa;    /* markaa: compressibility term, 1/kbar                     */
  double bb;    /* markbb: thermal expansion, 1/K                           */
  double cp;    /* markcp: heat capacity, J/kg/K                            */
  double kt;    /* markkt: conductivity, W/m/K                              */
  double kf;    /* markkf: conductivity T-term                              */
  double kp;    /* markkp: conductivity P-term                              */
} i2e_rock;

/* ---- run parameters (mode.t3c GENERAL/STOKES/HEAT blocks, loadmart.c:26-90,
 *      init.t3c header, in2mart.c:67-125) -----------------------------------*/
typedef struct i2e_params {
  int64_t xnumx, ynumy;       /* grid lines                                 */
  int64_t mnumx, mnumy;       /* markers per cell (lattice), for mardx      */
  double xsize, ysize;        /* box size, m                                */
  /* time-step controls (fl0stp, loadmart.c:44-46; maxtmstep in seconds)     */
  double maxxystep, maxtkstep, maxtmstep;
  /* STOKES_CONTINUITY_PARAMETERS                                            */
  double nubeg, nuend, nucontr, strmin;
  /* HEAT_TRANSFER_PARAMETERS                                                */
  double heatdif, frictyn, adiabyn;
  int32_t heatmod, densimod;
  int32_t outgrid, markmod;   /* markmod 1 = Euler, other = RK4 (markmart.c:74) */
  /* planet                                                                  */
  double gykoef;              /* GYKOEF: planet radius for gravity, m       */
  double tmp_ambient;         /* sticky-air temperature, K                  */
  double al2627_init;         /* 26Al/27Al (absolute, e.g. 5.25e-5)         */
  double fe6056_init;         /* 60Fe/56Fe (absolute)                       */
  double h0_scale;            /* multiplies the 26Al heating (1 = reference) */
  double longlived;           /* 1 = add 40K/235U/238U/232Th terms           */
  double por_init;            /* reference por_init (porosity before sinter) */
  double k_cutoff;            /* conductivity upper cut-off                  */
  double radmax_frac;         /* eta=1e22 beyond radmax_frac*xsize/2 (0.95)  */
  double eta_outer;           /* that viscosity (1e22)                       */
  /* partial-melting closure standing in for tdbasecalc's type switch
   * (markmart.c:2178-2182, meltpart1 linear branch markmart.c:2599-2613):
   * silicate types in melt_types[] switch to type+20 when the linear melt
   * fraction (T-Ts)/(Tl-Ts) is positive, and back when it is not; above
   * melt_switch the viscosity is melt_eta (clamped to [nubeg,nuend]), and
   * inside the melting interval cp gains melt_hlat*dX/dT (central difference
   * with 1 K steps, as markmart.c:2212-2219).                                */
  double melt_ts, melt_tl, melt_switch, melt_eta, melt_hlat;
  int32_t melt_on, porosity_on;
  int32_t ntypes;
  int32_t pad0;
  i2e_rock rock[I2E_MAXTYPES];
  uint8_t is_silicate[I2E_MAXTYPES];  /* radiogenic 26Al carrier (5,6,25,26) */
  uint8_t is_iron[I2E_MAXTYPES];      /* 60Fe carrier (7-10,17-19)            */
  uint8_t porous[I2E_MAXTYPES];       /* porosity closures (5,6)              */
  uint8_t meltable[I2E_MAXTYPES];     /* melt switch applies                  */
  /* soft-turbulence conductivity (markmart.c:503-610): once timesum >
   * core_form_time, molten silicate (types 25, 26) with melt fraction >= 0.4
   * and iron (7-10, 17-19) get the conductivity that makes the
   * soft-turbulence heat flux with the numerical viscosity equal to the one
   * with the fluid's (crystal mush, 1e2, 1e12 or 1e-2 Pa s).  The melt
   * fractions are the marker's own (linear silicate closure above; iron
   * between fe_ts and fe_tl, 0 if fe_tl is 0).                            */
  double core_form_time, fe_ts, fe_tl;
  int32_t soft_turb;
  /* magnetisation of solid silicates (markmart.c:651-744): markmg_old and
   * markmg_time of the host state follow the Curie temperatures of
   * magnetite (851.16 K) and hematite (948.16 K); dynamo = 1 while the
   * planetary dynamo runs (decided by core.c in the reference).           */
  int32_t magnet, dynamo;
  int32_t pad1;
  /* immobile rock types (markim): their markers are never advected
   * (markmart.c:27) and are left out of the temperature scatter
   * (markmart.c:905)                                                       */
  uint8_t immobile[I2E_MAXTYPES];
} i2e_params;

/* ---- host state: the reference globals as raw pointers ------------------ */
typedef struct i2e_host_state {
  double *gx, *gy;                    /* grid lines [xnumx], [ynumy]          */
  /* node fields, N = xnumx*ynumy each                                       */
  double *vx, *vy, *pr, *gp, *ga;
  double *exx, *eyy, *exy, *sxx, *syy, *sxy;
  double *nu, *nd, *ro, *tk, *cp, *kt, *ht, *et;
  /* markers                                                                 */
  int64_t marknum;
  float *markx, *marky, *markk, *markv, *markp, *markpor, *marktmax;
  int8_t *markt;
  /* boundary-condition tables in the reference layout (headmart.c:95-97):
   * bondm[4N] -> row of bondv/bondn (0 = no BC), bondn[bondnum][3] holds
   * 1-based unknown numbers of up to three neighbours (0 = none),
   * bondv[bondnum][4] = {const, k1, k2, k3}: x = const + sum k_i x_nbr_i.  */
  int64_t *bondm, *bondn;
  double *bondv;
  int64_t bondnum;
  /* scalars                                                                 */
  double timestep, timesum, gxkoef;
  /* magnetisation per marker (headmart.c markmg_old / markmg_time; NULL: not
   * kept by the caller, the device starts them at 0 and -1 Ma)              */
  int8_t *markmg_old;
  double *markmg_time;
} i2e_host_state;

/* ---- solver controls ---------------------------------------------------- */
enum { I2E_PREC_FP64 = 0, I2E_PREC_MIXED = 1 };
/* Pressure Schur-complement approximation of the Stokes preconditioner:
 * BFBt (two Galerkin-multigrid solves with B C^-1 B^T around one
 * velocity-block apply; C = diag(A), or with WBFBT its square root, i.e.
 * sqrt(viscosity) weights) or the viscosity-weighted pressure mass matrix. */
enum { I2E_SCHUR_BFBT = 0, I2E_SCHUR_MASS = 1, I2E_SCHUR_WBFBT = 2 };
/* Gram-Schmidt of the FGMRES solvers: lagged CGS2 (each basis vector's second
 * projection fused into the next iteration's pass: two passes over the basis
 * per iteration) or classical Gram-Schmidt repeated when the projection
 * removed most of the vector (Daniel-Gragg-Kaufman-Stewart).               */
enum { I2E_ORTH_LAGGED = 0, I2E_ORTH_DGKS = 1 };
/* Finest velocity level of the Stokes multigrid: matrix-free rows (in
 * difference form, which keeps FP32 accurate), or its probed 3x3 stencil held
 * in registers (fewer operations; in FP32 the rounded stencil needs ~6 % more
 * Krylov iterations), or -- mixed mode only -- the FP64 stencil applied to the
 * FP32 iterates with FP64 row sums (STENCIL64: the rows' iteration counts at
 * fewer operations), or -- mixed mode only -- the FP32 stencil in difference
 * form, sum s_k (u_k - u_centre) + (row sum) u_centre, which keeps the rows'
 * cancellation at half the bytes of STENCIL64 (DIFF32).  STENCIL64 and DIFF32
 * are STENCIL in FP64 mode.  AUTO: STENCIL in FP64 mode, DIFF32 in mixed.  */
enum { I2E_FINE_AUTO = 0, I2E_FINE_ROWS = 1, I2E_FINE_STENCIL = 2, I2E_FINE_STENCIL64 = 3, I2E_FINE_DIFF32 = 4 };
/* Initial guess of each Stokes solve: the last step's solution, or its
 * linear extrapolation in time from the last two solutions,
 * x_n + (dt / dt_prev) (x_n - x_{n-1}) (the first solve after i2e_upload
 * starts from the uploaded fields either way).  EXTRAP is the default:
 * 501^2, 6 steps: 75-77 Krylov iterations against 79-81 (mixed).          */
enum { I2E_GUESS_LAST = 0, I2E_GUESS_EXTRAP = 1 };

typedef struct i2e_solver_opts {
  int32_t precision;       /* I2E_PREC_FP64 or I2E_PREC_MIXED                 */
  int32_t restart;         /* FGMRES restart length (1..400, default 200)    */
  int32_t maxit;           /* max outer iterations per solve                 */
  int32_t nsmooth;         /* smoothing sweeps (pre = post): finest velocity
                              level, scalar (heat, gravity) multigrids      */
  double stokes_rtol;      /* ||b-Ax||/||b|| target for Stokes               */
  double heat_rtol;        /* same for heat                                  */
  double grav_rtol;        /* same for gravity                               */
  double schur_scale;      /* mass Schur: diagonal schur_scale*eta/mukoef     */
  double omega;            /* velocity smoother: colour-step relaxation in
                              (0, 2), 1 = Gauss-Seidel (default; 0.9, 1.1
                              and 1.2 each took 1-4 more iterations at 501^2) */
  int32_t coarse_max;      /* coarsest grid: cells per direction <= this     */
  int32_t schur;           /* Schur preconditioner: I2E_SCHUR_BFBT or _MASS  */
  int32_t graphs;          /* must be 0: graph replay of the Stokes
                              preconditioner is refused (I2E_ERR_UNSUPPORTED) */
  int32_t nsmooth_coarse;  /* velocity multigrid: sweeps below the finest level (0: nsmooth) */
  int32_t orth;            /* FGMRES orthogonalisation: I2E_ORTH_LAGGED or _DGKS */
  int32_t fine_stencil;    /* finest velocity level smoothed with I2E_FINE_*  */
  int32_t stokes_guess;    /* Stokes initial guess: I2E_GUESS_LAST or _EXTRAP */
  int32_t nsmooth_cell;    /* pressure-cell (BFBt) multigrid: sweeps on every
                              level (0: nsmooth).  Defaults 3 / 2 / 4: the
                              BFBt Poisson solves gain more from sweeps than
                              the velocity block (501^2: 82 vs 88 Krylov
                              iterations against 3 / 3 / 3, 7 % less time) */
  int32_t autotune;        /* 1: the tile shape of each coarse velocity level's
                              smoothing pass is timed among four candidates
                              at the first Stokes set-up and the fastest kept
                              (per level size, for the whole process; the
                              tiles change the block Gauss-Seidel partition,
                              so iteration counts can differ slightly);
                              0 (default): the fixed shapes                 */
  int32_t coarse_coop;     /* mixed precision: 1: the FP32 velocity levels of
                              at most 32 tiles of 32x16 points, down to the
                              two-level tail, in one cooperative launch per
                              V-cycle (the same arithmetic, bit for bit, as
                              one launch per pass); 0 (default): one launch
                              per pass.  On MI355X a cooperative launch
                              costs ~20 us before its first instruction and
                              each grid-wide barrier ~5 us, against ~4 us
                              per ordinary launch: 501^2 measured 63.1 ms
                              per step with it, 55.8 without              */
  int32_t stokes_direct;   /* 1: the Stokes FGMRES is preconditioned by an
                              LU factorisation of the Ruiz-equilibrated
                              rows R A C (band storage, partial pivoting
                              inside the band, FP64 whatever the
                              precision), refactored every solve: for
                              viscosity contrasts the multigrid cannot
                              precondition (melt against the clip).
                              Memory ~(3 kl + ku) 8 B per unknown, kl ~
                              3 ynumy (27 GB at 501^2).  2: the same, with
                              FGMRES on R A C y = R b, so that stokes_rtol
                              and stokes_relres refer to the equilibrated
                              residual ||R (b - A x)|| / ||R b|| (where the
                              melt makes ||b - A x|| / ||b|| <= 1e-8
                              unreachable in FP64); 0 (default): multigrid;
                              setting 0 frees the band storage            */
  int32_t pad_opts;
} i2e_solver_opts;

typedef struct i2e_stats {
  int32_t stokes_iters, heat_iters, grav_iters, pad;
  double stokes_relres, heat_relres, grav_relres;
  double numin, numax;      /* viscosity range after ronurecalc           */
  double ms_grav, ms_stokes, ms_heat, ms_markers;
} i2e_stats;

typedef struct i2e_ctx i2e_ctx;

/* Context: one HIP stream for all work, plus a second one on which the
 * Stokes set-up inverts the coarsest velocity matrix while the rest of the
 * set-up proceeds (the first stream waits for it before the solve, so
 * synchronising the first stream covers both); device copies of every
 * field.  max_markers bounds marker storage (markers can only leave through
 * the out-of-grid path).                                                    */
int i2e_create(const i2e_params *p, int64_t max_markers, i2e_ctx **out);
int i2e_destroy(i2e_ctx *c);
int i2e_set_solver(i2e_ctx *c, const i2e_solver_opts *o);
int i2e_default_solver(i2e_solver_opts *o);
/* Host <-> device.  upload copies everything the step reads (grid, nodes,
 * markers, BC tables, scalars); download copies everything it writes.       */
int i2e_upload(i2e_ctx *c, const i2e_host_state *s);
int i2e_download(i2e_ctx *c, i2e_host_state *s);

/* The reference's step functions, on device-resident state. */
int i2e_maxvelstep(i2e_ctx *c, double *timestep);        /* movemart.c:1225 */
int i2e_viterate(i2e_ctx *c, int n0);                    /* movemart.c:603  */
int i2e_titerate(i2e_ctx *c, int n0);                    /* heatmart.c:320  */
int i2e_movemark(i2e_ctx *c);                            /* markmart.c:3    */
int i2e_ronurecalc(i2e_ctx *c);                          /* markmart.c:192  */
/* One full step in i2mart.c:60-153 order (maxvelstep, viterate, maxvelstep,
 * titerate, movemark, ronurecalc, timesum += timestep).                     */
int i2e_step(i2e_ctx *c);
int i2e_get_scalars(i2e_ctx *c, double *timestep, double *timesum, double *gxkoef);
/* Per-cell marker-type histogram of the device-resident markers:
 * counts[(c1*(ynumy-1) + c2)*ntypes + t] = number of markers of type t in
 * cell (c1, c2), the cell being (m1serch, m2serch) of the marker position
 * (markmart.c:1051-1088) for markers strictly inside the box; host array of
 * (xnumx-1)*(ynumy-1)*ntypes ints.  The quantity the reference's scatter
 * (markmart.c:840-911) distributes per cell, exported for bit-exact checks. */
int i2e_type_counts(i2e_ctx *c, int ntypes, int32_t *counts);
int i2e_set_scalars(i2e_ctx *c, double timestep, double timesum);
int i2e_get_stats(i2e_ctx *c, i2e_stats *st);
/* Synchronise the context's stream. */
int i2e_sync(i2e_ctx *c);
/* Last error text (static storage, never NULL). */
const char *i2e_last_error(void);

/* ---- kernel-level entry points used by parity tests and the benchmark ---- */
/* y = K x with K the reference's assembled Stokes matrix (gausmat4 rows of
 * viterate, movemart.c:789-967); x, y are 3N interleaved device arrays.     */
int i2e_stokes_apply_dev(i2e_ctx *c, const double *x_dev, double *y_dev);
/* Same for the gravity rows (movemart.c:661-749) and heat rows
 * (heatmart.c:406-432, needs titerate's RHS state); N-length arrays.        */
int i2e_gravity_apply_dev(i2e_ctx *c, const double *x_dev, double *y_dev);
int i2e_heat_apply_dev(i2e_ctx *c, const double *x_dev, double *y_dev);
/* Diagnostics of the Stokes solver (no counterpart in the reference): the
 * right-hand side of viterate's rows, one application of the block
 * preconditioner z = P^-1 r (3N interleaved), and one velocity V-cycle on
 * f = [fx(N), fy(N)] -> u = [ux(N), uy(N)].  All device pointers.           */
int i2e_stokes_rhs_dev(i2e_ctx *c, double *b_dev);
int i2e_stokes_precond_dev(i2e_ctx *c, const double *r_dev, double *z_dev);
int i2e_vcycle_dev(i2e_ctx *c, const double *f_dev, double *u_dev);
/* Device pointers of the context's node/marker arrays, by reference name
 * ("vx", "tk", "markx", ...); NULL if unknown.  Marker arrays are in the
 * context's slot order, which the cell sort below permutes; "markorig"
 * gives the caller's index of every slot (int32; NULL while the slots are
 * in the caller's order).                                                   */
void *i2e_device_ptr(i2e_ctx *c, const char *name);
/* Cell sort of the device markers (no counterpart in the reference, which
 * keeps its markers in creation order): a counting sort on the cell key
 * m1serch*(ynumy-1) + m2serch, markers outside the box last, so that the
 * marker kernels read whole cells per wavefront.  It changes no result
 * beyond the summation order of the scatter; i2e_download returns markers
 * in the caller's order and i2e_upload restores it.  i2e_step sorts every
 * `every` steps (default 10; 0 = never; suspended while i2e_record runs);
 * i2e_sort_markers sorts now.  Extra device memory: 33 B per marker slot. */
int i2e_set_marker_sort(i2e_ctx *c, int every);
/* The per-step hydration and body-statistics sweep of impact()
 * (impact.c:3022-3228) on the device-resident state.  Nodes: maxtk = max of
 * tmp_ambient and every tk, meantk = mean tk over the nodes whose tk differs
 * from tmp_ambient.  Markers, when `hydrate` is nonzero (the reference's
 * growth_model == 0 && por_init <= 0.001): type 5 becomes 6 once marktmax
 * reaches 273.15 K (below 1223.15 K) and 6 becomes 5 at 1223.15 K; then the
 * silicate markers (types 5, 6, 25, 26) are counted by their current type
 * and by marktmax against the classification thresholds, with the body's
 * maximum and mean marker temperature and the markers hotter than 1.1 maxtk.
 * t_sol and t_liq are impact()'s peridotite solidus and liquidus (1393.811
 * or 1416.193 K and 1973 K, impact.c:188-198).  Fractions are counts over
 * count_tot (0/0 when no marker counts, as in the reference).            */
typedef struct i2e_body_stats {
  double sol_frac, liq_frac, hydrous_frac, primitive_frac, n2co2_frac, cocl_frac, h2o_frac;
  double phyllo1_frac, phyllo2_frac, phyllo3_frac, phyllo4_frac, perco_frac, melt1_frac, melt2_frac;
  double maxtk, t_max_body, meantk, t_mean_body;
  int64_t count_toohot, count_tot, count_body_nodes;
} i2e_body_stats;
int i2e_hydration(i2e_ctx *c, double t_sol, double t_liq, int hydrate, i2e_body_stats *out);

/* The tile shapes the autotuner chose (autotune = 1), one entry per coarse
 * velocity level smoothed by tiles, finest first: the level's nodes along x
 * in levels[], the tile's width (along y) and height in sx[], sy[] (arrays
 * of nmax, any may be NULL); *n = the number of such levels.  No
 * counterpart in the reference (its solves are PARDISO's).                 */
int i2e_tuned_shapes(i2e_ctx *c, int nmax, int32_t *levels, int32_t *sx, int32_t *sy, int *n);
int i2e_sort_markers(i2e_ctx *c);
/* Advection kernel alone, timed with HIP events on the context's stream:
 * runs movemark's marker loop `reps` times from the same start positions
 * and returns the mean kernel milliseconds.                                 */
int i2e_time_kernels(i2e_ctx *c, int reps, double *ms_advect, double *ms_p2g,
                     double *ms_stokes_apply);

/* Ensemble: nsteps of i2e_step on each of n independent bodies, bodies
 * spread over nthreads host threads (0: one per body) so that their
 * launches overlap on the GPU.  status[k] (may be NULL) receives body k's
 * result; the return value is the first failure.  No counterpart in the
 * reference, which runs one body per process (i2mart.c:60-164).           */
int i2e_step_batch(i2e_ctx *const *ctxs, int n, int nsteps, int nthreads, int *status);

/* Residual diagnostics of viterate (movemart.c:1049-1176, printed there
 * when printmod is set): RMS of the x/y-Stokes errors in their error form
 * (stresses of the current fields, divided by the mean nodal viscosity) over
 * (max velocity range / min grid step^2), of the continuity errors over
 * (range / step) and of the velocity BC rows over the range; row counts and
 * the velocity, pressure and viscosity ranges the reference prints.  Works
 * on the context's current fields (call after i2e_viterate).              */
typedef struct i2e_resid {
  double stokes, cont, bond;
  double stoknum, contnum, bondnum;
  double vxmin, vxmax, vymin, vymax, prmin, prmax, numin, numax;
} i2e_resid;
int i2e_stokes_residuals(i2e_ctx *c, i2e_resid *out);
/* Bandwidths of the last banded factorisation (stokes_direct = 1): kl
 * subdiagonals and ku superdiagonals of the assembled Stokes rows in the
 * reference's unknown order (gausmat4's system, gausmart.c:174-404), and
 * the device bytes of its band storage.                                     */
int i2e_band_info(i2e_ctx *c, int32_t *kl, int32_t *ku, double *bytes);

/* Per-launch HIP-event timing of one kernel on the context's stream (no
 * counterpart in the reference), for roofline figures taken inside a normal
 * run: i2e_kernel_timer selects the kernel ("advect" = movemark's marker
 * pass, "p2g" = ronurecalc's marker scatter, "gsdot" / "gsaxpy" = the
 * Gram-Schmidt projection / update passes of the Stokes FGMRES,
 * "vsmooth0" / "vsmoothc" = one colour step of the velocity smoother on the
 * finest / a coarse multigrid level, "stokes_apply", "precond" = one
 * whole Stokes preconditioner application (bytes not counted); NULL or ""
 * stops timing), reserves events for up to max_launches timed launches,
 * times every `every`-th launch of the kernel (1: all; event pairs add
 * latency to the stream) and clears the counters; i2e_kernel_timer_read
 * returns the launches timed, their summed milliseconds and summed
 * algorithmic HBM bytes.                                                    */
int i2e_kernel_timer(i2e_ctx *c, const char *kernel, int max_launches, int every);
int i2e_kernel_timer_read(i2e_ctx *c, int64_t *launches, double *ms_total,
                          double *bytes_total);
/* The representation bytes of the same launches: what this implementation
 * reads and writes to do the work (e.g. the explicit Galerkin stencils of
 * the coarse smoothing passes, 4 x 106 B per point, against the ~40 B per
 * point and sweep of the matrix-free algorithm the timer's algorithmic
 * bytes count).  Timer names: also "heatmark" = the marker temperature
 * update's three launches (titerate, heatmart.c:515-689) as one unit; "p2g"
 * covers the scatter and the node normalisation launch.                   */
int i2e_kernel_timer_repr(i2e_ctx *c, double *repr_total);

/* Replace the run parameters of a context (the reference reads them once
 * from mode.t3c/init.t3c, loadmart.c:26-90, into globals).  Grid sizes and
 * markers per cell must match the context's; the change applies from the
 * next call on.                                                             */
int i2e_set_params(i2e_ctx *c, const i2e_params *p);

/* ---- adjoint of a peak-temperature objective (no counterpart in the
 * reference: a sensitivity study of the same time step) --------------------
 *
 * J = sum_i m_i marktmax_i / sum_i m_i over silicate markers inside the
 * grid, m_i = markro(type) (1 - markpor_i).  Parameters: h0_scale (the 26Al
 * heating multiplier, markmart.c:758-770) and phi0, a uniform shift of every
 * marker porosity that is positive at the start.
 *
 * i2e_record runs i2e_ronurecalc and then nsteps of i2e_step from the
 * uploaded state, keeping in device memory the marker positions and
 * temperatures after every ronurecalc and the node fields of every heat
 * solve.  i2e_replay runs the thermal part of the same steps (ronurecalc,
 * titerate) from the uploaded state along the recorded positions and time
 * steps, without Stokes, gravity or advection.  i2e_adjoint_peak_t, called
 * right after i2e_record, returns J and its gradient through that thermal
 * chain (titerate heatmart.c:320-695, ronurecalc markmart.c:192-1019) with
 * positions and time steps held at the recorded ones: the exact gradient of
 * what i2e_replay computes, and of the full step whenever the flow does not
 * depend on the parameters.  Requires melt_on = porosity_on = 0 and
 * temperature/pressure-independent density and conductivity (markaa =
 * markbb = markkf = markkp = 0); replay also requires frictyn above the run
 * time (no shear heating).  Device memory: 12 B per marker and 48 B per node
 * per recorded step.                                                        */
typedef struct i2e_adjoint_out {
  double J;
  double dJ_dh0;           /* dJ / d h0_scale                                */
  double dJ_dphi0;         /* dJ / d phi0                                    */
  double heat_relres_max;  /* largest residual of the transposed heat solves */
  int32_t steps, heat_iters;
} i2e_adjoint_out;
int i2e_record(i2e_ctx *c, int nsteps);
int i2e_replay(i2e_ctx *c, int nsteps);
int i2e_peak_t(i2e_ctx *c, double *J);
int i2e_adjoint_peak_t(i2e_ctx *c, i2e_adjoint_out *out);
/* y = A^T x for the heat rows of i2e_heat_apply_dev (device pointers).     */
int i2e_heat_apply_t_dev(i2e_ctx *c, const double *x_dev, double *y_dev);

/* The same objective J through the whole recorded step (no counterpart in
 * the reference): i2e_adjoint_peak_t_coupled, called right after
 * i2e_record, returns J and its gradient with respect to h0_scale, phi0 and
 * s, a common factor on every rock's activation energy markdh (at s = 1),
 * with the flow differentiated too: the marker positions move with the
 * Stokes velocities (RK4 or Euler, markmart.c:74-115, reversed from the
 * positions recorded at every step), the velocities solve the Stokes rows
 * (movemart.c:4-426; transposed solve with viterate's preconditioner) whose
 * viscosities, density and gravity potential (movemart.c:622-785,
 * transposed gravity solve) come from ronurecalc's scatter at those
 * positions (markmart.c:840-1009, weights differentiated in the positions),
 * and the viscosities from viscalc's Arrhenius branch of the marker
 * temperatures (markmart.c:1959-1977).  Time steps are held: records whose
 * dt maxvelstep or titerate's heating-rate limit shortened are refused
 * (raise maxxystep and maxtkstep), as are non-uniform grids, shear heating
 * and the cases i2e_adjoint_peak_t and the flow adjoint refuse.  Device
 * memory: the record's 12 B per marker and 48 B per node per step plus 4 B
 * per marker and 56 B per node per step.  The call leaves the context's
 * node fields at those of the first recorded step (upload before stepping
 * again).                                                                   */
typedef struct i2e_coupled_adjoint_out {
  double J;
  double dJ_dh0, dJ_dphi0, dJ_dEscale;
  double heat_relres_max, stokes_relres_max, grav_relres_max;
  int64_t heat_iters, stokes_iters, grav_iters;
  int32_t steps, pad;
} i2e_coupled_adjoint_out;
int i2e_adjoint_peak_t_coupled(i2e_ctx *c, i2e_coupled_adjoint_out *out);
/* dJ/d(initial marker positions) from the last i2e_adjoint_peak_t_coupled
 * of the current record: host arrays of marknum doubles, caller's order.    */
int i2e_adjoint_marker_positions(i2e_ctx *c, double *dJ_dx, double *dJ_dy);

/* ---- flow adjoint: one Stokes solve and the advection that reads it ------
 * No counterpart in the reference: a sensitivity of the iron sinking rate
 *   J = sum_i m_i (v_i . r_i) / sum_i m_i
 * over the iron markers inside the grid (markk > 0, is_iron, mobile), m_i =
 * markro of the type, r_i the unit vector from the box centre, v_i the
 * velocity movemark gives marker i (RK4, or Euler for markmod 1;
 * markmart.c:74-97) from the node velocities of the last viterate with the
 * context's current timestep.  i2e_sink_rate returns J.
 * i2e_adjoint_sink_rate, called after i2e_viterate (and i2e_maxvelstep) and
 * before anything that changes the markers or node fields, returns J and
 * dJ/ds, s a common factor on every rock's activation energy markdh at
 * s = 1, through: the reversed RK4 interpolation (allinterv, markmart.c:
 * 1378-1453), the transposed Stokes solve K^T mu = dJ/dx (rows of
 * movemart.c:4-426) by FGMRES with viterate's preconditioner, the viscosity
 * scatter and node normalisation of ronurecalc (markmart.c:840-1009) and
 * viscalc's Arrhenius branch (markmart.c:1855-2063).  Marker positions and
 * temperatures, density, gravity and the time steps are held.  Refuses melt
 * switching, activation volumes, negative activation energies, power-law
 * rheologies (I2E_ERR_UNSUPPORTED) and node viscosities set by the contrast
 * limit nucontr.                                                            */
typedef struct i2e_flow_adjoint_out {
  double J;              /* m/s (positive outward)                           */
  double dJ_dEscale;     /* dJ / ds                                          */
  double stokes_relres;  /* residual of the transposed Stokes solve          */
  int64_t stokes_iters;  /* its FGMRES iterations                           */
} i2e_flow_adjoint_out;
int i2e_sink_rate(i2e_ctx *c, double *J);
int i2e_adjoint_sink_rate(i2e_ctx *c, i2e_flow_adjoint_out *out);
/* y = K^T x for the Stokes rows of i2e_stokes_apply_dev (device pointers).  */
int i2e_stokes_apply_t_dev(i2e_ctx *c, const double *x_dev, double *y_dev);

/* ---- post-processing: the visualisation grid of i2jmart's saver ---------
 * i2jmart (i2jmart.c, loadjmart.c) turns a .prn file into a regular grid of
 * one field for plotting (Amira).  i2e_visual computes that grid on the
 * device from the context's state, as saver() does (loadjmart.c:619-1197,
 * one x-chunk): markers inside the window (every markstp-th, markt < 50) add
 * their value and bilinear weight to the four surrounding pixels in float32,
 * in marker order; a pixel with weight divides, one without interpolates the
 * node fields at its position (allinter1, loadjmart.c:1327-1520).  type
 * (amir.t3c's letter): 1 T (K), 2 log10 viscosity, 4 log10 shear heating,
 * 5 rock types (the type of largest summed weight per pixel, in type order,
 * empty pixels filled from neighbours within gridmod pixels), 6 density,
 * 7 conductivity, 8 heat generation (the rock table's markht is not carried:
 * shear heating only on markers, plus the node ht field in empty pixels),
 * 12 vorticity / strain-rate ratio (spncalc, loadjmart.c:1523-1617).  Types
 * 3 (the reference leaves the marker value unset), 9-11 (seismic velocities,
 * never computed there) and densimod 2 are refused (I2E_ERR_UNSUPPORTED).
 * xbeg..yend are fractions of the box (clamped to [0, 1]); xresol, yresol
 * >= 2, xresol * yresol * 2 <= 40000000 (the reference's val0[]).  Outputs,
 * each xresol * yresol long, pixel (m1, m2) at m1 * yresol + m2: val (types
 * != 5) or rock (type 5; -1 where none).  melt (optional, marknum doubles):
 * the melt fraction saver() computes per visited marker for markersave()
 * (0 elsewhere).  Markers with markk < 1 take the interpolated node T and
 * raise their marktmax as saver() does; the context's markers are left as
 * they are.                                                                 */
typedef struct i2e_visual_spec {
  int32_t type, gridmod;
  int32_t xresol, yresol;
  double xbeg, xend, ybeg, yend;
} i2e_visual_spec;
int i2e_visual(i2e_ctx *c, const i2e_visual_spec *spec, float *val, int32_t *rock, double *melt);
/* The text file saver() writes (loadjmart.c:681, 1130-1165): the time in
 * years and the resolution, then every pixel in m1-outer order as
 * " % 5.4e", or for type 5 the rock types with runs of more than three equal
 * values along y written as " -2 count type".                                */
int i2e_visual_write(const char *path, const i2e_visual_spec *spec, double timesum, const float *val,
                     const int32_t *rock);
/* markersave (loadjmart.c:3-53): marker_melt.txt, marker_comp.txt,
 * marker_magnet.txt and marker_magnet_time.txt in directory dir, one line
 * per marker (x / xsize, y / ysize, value), marknum + 1 lines as the
 * reference writes them (the last marker record is zero).  melt may be NULL
 * (zeros).                                                                  */
int i2e_markersave(const char *dir, const i2e_params *p, const i2e_host_state *s, const double *melt);

/* grain() (grain.c:4-156), called by main() after each output cycle
 * (i2mart.c:178): olivine grain growth per marker, d <- (d^n + k0
 * exp(-E/(R T)) dt)^(1/n) with the context's last timestep, for olivine in
 * molten FeS (types 5, 25) after the first impact and pure olivine (6, 26)
 * after the impact (growth_model 1) or the model start (2); with growth
 * models 1 and 2 and t + dt in [impact, impact + 0.45 Ma], markers in the
 * mantle sector 100-200 km from the centre at 45 +- 10 degrees become
 * pallasite (types 6 -> 5, 26 -> 25) with the initial grain size.  markgr:
 * the marknum grain sizes (markgr of headmart.c:124, micrometres) in the
 * caller's order, updated in place; marker types change in the context.
 * impact_time_ma: impact_time[1] of the impact history (Ma after CAIs).    */
typedef struct i2e_grain_params {
  int32_t growth_model, pad;
  double gr_init, start_time_ma, impact_time_ma;
} i2e_grain_params;
int i2e_grain(i2e_ctx *c, const i2e_grain_params *gp, float *markgr);

/* ---- binary .prn checkpoint (saver/loader of loadmart.c) -----------------
 * The byte layout of the reference's binary output (saver, loadmart.c:671-809;
 * loader, loadmart.c:391-540; read by support_files/plot2d.py:157-283), for a
 * host state: node fields, grid lines and BC values are stored as float32,
 * as there.  Fields this build does not model are written as zero and
 * skipped on reading; like the reference, the file holds marknum + 1 marker
 * records (the last one zero here).  i2e_prn_read_header gives the sizes to
 * allocate; i2e_prn_read then fills a state of exactly those sizes and the
 * params the file carries (grid, rock table, 26Al/60Fe, ambient T, GYKOEF),
 * leaving the others (time-step and solver controls, which the reference
 * reads from mode.t3c) as the caller set them.  gp and ga are not in the
 * file and come back zero.                                                  */
typedef struct i2e_prn_header {
  int64_t xnumx, ynumy, mnumx, mnumy, marknum, bondnum;
  int32_t rocknum, n0;
  double xsize, ysize, por_init, gxkoef, gykoef, tmp_ambient;
  double al2627_init, fe6056_init;  /* absolute ratios (the file scales them) */
  double timesum;                   /* seconds (the file holds years)         */
} i2e_prn_header;
int i2e_prn_write(const char *path, const i2e_params *p, const i2e_host_state *s, int n0);
int i2e_prn_read_header(const char *path, i2e_prn_header *h);
int i2e_prn_read(const char *path, i2e_params *p, i2e_host_state *s, int *n0);

/* maxvelstep (movemart.c:1225-1262) on caller-owned device arrays, for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops.  No counterpart in the reference.  vx, vy: N = xnumx
 * * ynumy FP64 node fields (node m = m1*ynumy + m2); *timestep_dev (one FP64
 * on the device) holds the current time step and receives
 * min(dt, maxxystep*dx/max|vx|, maxxystep*dy/max|vy|), dx = xsize/(xnumx-1),
 * dy = ysize/(ynumy-1), bit-identical to i2e_maxvelstep.  Launches three
 * kernels on `stream` (a hipStream_t; NULL = the null stream) and returns
 * without waiting; allocates nothing and calls no hipMalloc/hipFree/memset.
 * Its scratch (the two maxima, 32 bytes) is a device array of the library,
 * one per GPU, kept until the process ends: calls that overlap on one GPU
 * race on it and must be ordered by the caller (one stream, or events).
 * Grids above 16385 per direction or 2^28 nodes: I2E_ERR_CAPACITY.       */
int i2e_maxvelstep_noalloc(const double *vx_dev, const double *vy_dev, int64_t xnumx, int64_t ynumy,
                           double xsize, double ysize, double maxxystep, double *timestep_dev, void *stream);

/* ---- initial state from an init.t3c file (in2mart.c:23-1453) -------------
 * The state a reference run starts from, built with in2mart's semantics:
 * grid lines (with X/Y grid-line definitions), the marker lattice and its
 * rand() jitter (the C library's sequence without srand, as the reference
 * has it), the rock table, the boundary-condition tables (regular boxes,
 * irregular rectangles and circular sectors for Vx, Vy, P, T; continuity
 * coefficients 'c'; cell marker sets), rock bodies on the markers,
 * the initial temperature field (box, exponential, circular, cooling-age,
 * transition and impact distributions), porosity and the marker
 * temperatures interpolated from it.  Not handled (I2E_ERR_UNSUPPORTED):
 * markers loaded from a file and rock bodies flagged immobile.
 * i2e_init_read gives the sizes; i2e_init_generate fills a state of exactly
 * those sizes and the params the file holds (grid, rock table with its
 * immobile types, 26Al/60Fe, ambient T, GYKOEF, por_init).  Node fields are FP64 as in2mart holds them (its binary save
 * rounds them to float32: write and read back a .prn for that).          */
typedef struct i2e_init_info {
  int64_t xnumx, ynumy, mnumx, mnumy, marknum, bondnum;
  int32_t rocknum, nonstab;
  double timesum;                   /* seconds                                */
} i2e_init_info;
int i2e_init_read(const char *init_path, i2e_init_info *info);
int i2e_init_generate(const char *init_path, i2e_params *p, i2e_host_state *s);
/* The run controls of a mode.t3c file (loadconf, loadmart.c:2-105) into p:
 * from output-file line `file_index` (0-based) the time-step limits
 * maxxystep, maxtkstep, maxtmstep (years in the file, seconds in p) and its
 * cycle count (*cyc0max, may be NULL); markmod, outgrid, densimod; nubeg,
 * nuend, nucontr, strmin; heatmod, heatdif, frictyn and adiabyn (years in
 * the file).  Other fields of p are left as they are.                     */
int i2e_mode_read(const char *mode_path, int file_index, i2e_params *p, int *cyc0max);

/* Device buffers for callers without a device allocator of their own (no
 * counterpart in the reference): allocation, release and blocking copies,
 * all through the same HIP runtime as the library itself.                   */
int i2e_dev_alloc(size_t bytes, void **ptr);
int i2e_dev_free(void *ptr);
int i2e_copy_to_dev(void *dst_dev, const void *src_host, size_t bytes);
int i2e_copy_to_host(void *dst_host, const void *src_dev, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (k_dev_copy: one read and
 * one write per 16-byte element; bytes and both addresses multiples of 16),
 * the known-size reference for scaling rocprof's FETCH_SIZE / WRITE_SIZE.   */
int i2e_dev_copy(void *dst_dev, const void *src_dev, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* I2E_H */
