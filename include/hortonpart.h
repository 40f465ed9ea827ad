/*
 * hortonpart.h -- C ABI of the MI355X stockholder-partitioning library.
 *
 * The reference (horton-part 1.1.7) is a pure-Python package with no FFI; its
 * boundary for this path is the WPart scheme-class API.  The entry points
 * below are what that API's host classes (horton_part_amd.mbis / .gisa /
 * .alisa) bind through ctypes.  Each one names the reference routine whose
 * work it takes over (paths relative to the reference's src/horton_part).
 *
 * Conventions
 *  - plain C types only: pointers to host memory, sizes, doubles, ints;
 *  - every array is C-contiguous; the library copies inputs to the GPU;
 *  - the caller owns every output buffer;
 *  - every function returns 0 on success and a negative HP_E_* code on
 *    failure; hp_last_error() then describes the failure (thread-local);
 *  - one hp_part object per host thread at a time (not re-entrant).
 */
#ifndef HORTONPART_H
#define HORTONPART_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HP_ABI_VERSION 3

/* error codes (negative returns) */
#define HP_E_ARG -1      /* invalid argument or shape */
#define HP_E_HIP -2      /* HIP runtime error (no device, launch failure...) */
#define HP_E_SOLVER -3   /* a per-atom solver raised (see hp_part_status) */
#define HP_E_UNSUPP -4   /* combination not implemented on the device */

/* partitioning schemes */
#define HP_SCHEME_MBIS 0 /* mbis.py:206-362 */
#define HP_SCHEME_GISA 1 /* gisa.py:109-345 */
#define HP_SCHEME_LISA 2 /* alisa.py:1155-1335 */
#define HP_SCHEME_NLIS 3 /* nlis.py:201-367 and gmbis.py:75-156: (N, S, n) per shell */
#define HP_SCHEME_ISA 4  /* isa.py:37-122: spline pro-atoms through the spherical averages */
#define HP_SCHEME_HI 5   /* hirshfeld_i.py:36-176: interpolated pro-atom database records */
#define HP_SCHEME_HIRSHFELD 6 /* hirshfeld.py:127-193: fixed neutral records, one weights pass */

/* inner solvers */
#define HP_SOLVER_MBIS 0     /* mbis.py:81-163 opt_mbis_propars */
#define HP_SOLVER_QP 1       /* gisa.py:348-421 opt_propars_qp_interface */
#define HP_SOLVER_SC 2       /* alisa.py:193-291 solver_sc */
#define HP_SOLVER_NEWTON 3   /* alisa.py:543-610 solver_newton (exact) */
#define HP_SOLVER_MNEWTON 4  /* alisa.py:613-680 solver_m_newton (line search) */
#define HP_SOLVER_NLIS 5     /* nlis.py:99-195 opt_nlis_propars */

/* arithmetic of the grid passes */
#define HP_PREC_FP64 0  /* everything in FP64 */
#define HP_PREC_MIXED 1 /* FP32 exp/promolecule/ratio, FP64 sums and solves */

/* per-atom status bits (hp_part_status) */
#define HP_ST_INNER_NOT_CONVERGED 1 /* mbis.py:162 / alisa.py:289 warning */
#define HP_ST_POP_MISMATCH 2        /* mbis.py:157 / utils.py:442-456 warning */
#define HP_ST_NEG_DENSITY 4         /* utils.py:512-520 RuntimeError */
#define HP_ST_LINESEARCH 8          /* alisa.py:877 RuntimeError */
#define HP_ST_NEWTON_MAXITER 16     /* alisa.py:938 RuntimeError */
#define HP_ST_QP_FAILED 32          /* quadprog failure (gisa.py:409) */
#define HP_ST_NEG_PARS 64           /* utils.py:482-489 warning */
#define HP_ST_NOT_MONOTONIC 128     /* utils.py:546-553 RuntimeError */
#define HP_ST_SINGULAR 256          /* alisa.py:924 scipy LinAlgError (singular Hessian) */
#define HP_ST_NONFINITE 512         /* alisa.py:924 scipy ValueError (inf/NaN Hessian) */
#define HP_ST_NORECORD 1024         /* hirshfeld_i.py:124-157: no database record for the charge
                                       (KeyError) or a pseudo population <= 0 (ValueError) */

/* A molecule on a grid of atom-centred product grids.  The molecular grid is
 * the concatenation of the atom grids (atom-major, radial-shell-major), as
 * qc-grid's MolGrid builds it (core/base.py:490-499, 619-628). */
typedef struct {
  int natom;
  const double *coordinates;    /* natom*3, bohr */
  const int64_t *numbers;       /* natom */
  const double *pseudo_numbers; /* natom */
  int64_t npoint;
  const double *points;      /* npoint*3 */
  const double *density;     /* npoint: molecular density rho */
  const double *mol_weights; /* npoint: atom-grid weight x Becke weight */
  const double *atom_weights;/* npoint: atom-grid weight (moments) */
  const double *sph_weights; /* npoint: angular weight within its shell (sum 1) */
  int nshell;                /* total number of radial shells */
  const int32_t *shell_atom; /* nshell: owning atom */
  const int64_t *shell_offsets; /* nshell+1: point offsets */
  const double *shell_radius;   /* nshell: r_i */
  const double *shell_rweight;  /* nshell: radial quadrature weight w_i */
} hp_molecule_desc;

/* The scheme: per-atom pro-atom functions and solver options. */
typedef struct {
  int scheme;    /* HP_SCHEME_* */
  int solver;    /* HP_SOLVER_* */
  int grid_type; /* 1: atom grids (default); 2: molecular grid; 3: molecular
                    grid only (the loop of 2; moments on the molecular grid) */
  const int32_t *fn_offsets; /* natom+1: functions of atom a are [off[a], off[a+1]) */
  const double *fn_alpha;    /* GISA/LISA: exponent of each function */
  const double *fn_order;    /* GISA/LISA/NLIS: power n of r in exp(-alpha r^n)
                                (NLIS: the n of the parameters, fixed) */
  const double *fn_norm;     /* GISA/LISA: n alpha^(3/n) / (4 pi Gamma(3/n)) */
  const double *qp_matrix;   /* GISA: concatenated K_a x K_a blocks of P */
  double threshold;          /* outer threshold (iterstock.py:99) */
  double inner_threshold;    /* already min(inner, threshold) (iterstock.py:101) */
  int maxiter;               /* outer cap (iterstock.py:187) */
  int inner_maxiter;         /* 2000 (MBIS), 1e5 (sc), 10000 (newton) */
  double density_cutoff;     /* sick-point cutoff (constants.yaml:27) */
  double negative_cutoff;    /* constants.yaml:28 */
  double population_cutoff;  /* constants.yaml:29 */
  double cutoff_radius;      /* <= 0: every pro-atom on every point */
  double newton_tau;         /* m-newton line-search start (alisa.py:772) */
  int newton_linspace;       /* m-newton line-search size (alisa.py:773) */
  int check_mono;            /* alisa.py:774 */
  double mixed_switch;       /* HP_PREC_MIXED: leave the FP32 stage once the
                                change is below max(1.01 threshold, mixed_switch)
                                (grid_type 2: max(threshold, mixed_switch), where
                                <= 0 selects 1e-5); the Python classes pass 1e-5 */
  int mixed_refresh;         /* HP_PREC_MIXED, grid_type 1: every this many
                                FP32-stage iterations one runs in FP64 and
                                refreshes the per-shell correction FP64 - FP32
                                that the FP32 iterations add to their spherical
                                averages; <= 0: no correction */
  int full_sum;              /* MBIS, GISA: 0 (default) lets the weights pass
                                skip, per radial shell, every atom whose
                                pro-atom is bounded there by 2^-64 (FP32 stage:
                                2^-40) of the promolecule's lower bound (the
                                shell atom's pro-atom at the shell's largest
                                radius plus the 1e-100 floors): the skipped
                                terms move the promolecule by less than one
                                rounding.  1: every atom on every point, as the
                                reference evaluates them.  LISA always sums
                                every atom (its coefficients may be negative). */
  /* Tabulated pro-atoms (HP_SCHEME_ISA, HP_SCHEME_HI, HP_SCHEME_HIRSHFELD;
   * solver and the basis arrays above are ignored, fn_offsets may be NULL).
   * Atom a's pro-atom is a piecewise cubic over the ascending radial knots
   * tab_knots[tab_offsets[a] .. tab_offsets[a+1]) (at least 2; at most 512),
   * whose radial quadrature weights tab_rweights enter the change
   * (iterstock.py:32-45).  ISA: the knots are the atom's radial shells in
   * order (at least 4) and the tables are built on the device from the
   * parameters; HI, HIRSHFELD: each atom's element grid from the database. */
  const int32_t *tab_offsets; /* natom+1 */
  const double *tab_knots;
  const double *tab_rweights;
  /* HP_SCHEME_HI: the database records of each atom's element,
   * [rec_offsets[a], rec_offsets[a+1]) for atom a, record j of integer charge
   * rec_charges[j] (ProAtomRecord, core/proatomdb.py:34-210).  Record values
   * are stored in record order, each on its atom's knots: densities rec_rho
   * (n_a per record) and the pieces of its spline rec_coefs (4 n_a per
   * record: c0 d^3 + c1 d^2 + c2 d + c3 at d = r - knot_i for i < n_a - 1, the
   * last row unused; ProAtomDB.get_spline, proatomdb.py:378-388).
   * HP_SCHEME_HIRSHFELD: exactly one record per atom, its fixed pro-atom
   * (get_proatom_spline of the neutral record, core/stockholder.py:227-264). */
  const int32_t *rec_offsets; /* natom+1 */
  const int32_t *rec_charges;
  const double *rec_rho;
  const double *rec_coefs;
} hp_scheme_desc;

typedef struct hp_part hp_part;

/* Library ABI version (HP_ABI_VERSION). */
int hp_abi_version(void);

/* Text of the last error on this thread ("" if none). */
const char *hp_last_error(void);

/* Number of visible GPUs. */
int hp_device_count(int *count);

/* Upload a molecule and its scheme and allocate the device state.
 * Replaces the construction work of WPart.__init__ (core/base.py:416-488),
 * calc_radial_distances (core/base.py:630-635; distances are recomputed on
 * the fly instead of stored) and the basis set-up of evaluate_basis_functions
 * (gisa.py:91-106). */
int hp_part_create(int device, const hp_molecule_desc *mol, const hp_scheme_desc *scheme, hp_part **out);

/* Release all device memory of a partitioning. */
int hp_part_destroy(hp_part *part);

/* Set the current pro-atom parameters (initial values: mbis.py:291-305,
 * gisa.py:68-88, nlis.py:330-346, gmbis.py:138-152).  Length: 2*sum(K_a)
 * for MBIS, 3*sum(K_a) for NLIS (whose n must equal fn_order), sum(K_a)
 * for GISA/LISA; ISA: the pro-atom values at the knots (sum n_a, initially
 * zero, isa.py:93-100); HI: the charges (natom, initially zero,
 * hirshfeld_i.py:160-164); HIRSHFELD: natom values that are not used. */
int hp_part_set_propars(hp_part *part, const double *propars, int64_t n);

/* One update_at_weights pass (core/stockholder.py:352-384) at the current
 * parameters, in FP64, for partitionings whose per-atom solver runs in the
 * caller: the solver callables of GaussianISAWPart._opt_propars
 * (gisa.py:320-341) and LinearISAWPart._opt_propars (alisa.py:1304-1335).
 * rhobar (grid_type 1, nshell values; may be NULL) receives the spherical
 * averages of the AIM densities at the radial knots, the
 * "spherical_average_{a}" entries of gisa.py:281-299; entropy (may be NULL)
 * the entropy of the promolecule (stockholder.py:145-151).  Afterwards
 * hp_part_at_weights and hp_part_promoldens return this pass's weights
 * and promolecule until the next pass or run. */
int hp_part_weights(hp_part *part, double *rhobar, double *entropy);

/* Run the outer fixed-point loop of AbstractISAWPart.do_partitioning
 * (core/iterstock.py:159-193): per iteration update_at_weights
 * (core/stockholder.py:352-384), the per-atom updates (mbis.py:167-203,
 * gisa.py:251-318, alisa.py:1304-1335), the entropy (stockholder.py:145-151)
 * and compute_change (iterstock.py:32-45).  precision: HP_PREC_*.  With
 * HP_PREC_MIXED the loop runs in mixed precision until the change is below
 * max(1.01 threshold, mixed_switch) (FP32 bulk with the defect correction
 * of mixed_refresh); the iteration that reaches it is discarded and redone in
 * FP64, and FP64 iterations (at least n_polish) follow until the change is
 * below threshold; maxiter caps the total number of iterations unless the
 * polish needs more.  Returns the number of
 * outer iterations through *niter.  HP_E_SOLVER if a per-atom solver raised;
 * the loop stops at that iteration.  ISA and HI iterate in FP64 (their
 * per-atom updates are closed forms: isa.py:102-122, hirshfeld_i.py:166-172);
 * HIRSHFELD does not iterate (HP_E_UNSUPP: use hp_part_weights). */
int hp_part_run(hp_part *part, int precision, int n_polish, int *niter);

/* Copy results to host buffers (any pointer may be NULL):
 * propars (P), charges (natom), last change and entropy. */
int hp_part_results(hp_part *part, double *propars, double *charges, double *change, double *entropy);

/* Histories (iterstock.py:51-54): niter rows each.  history_propars niter*P,
 * history_charges niter*natom, history_entropies niter, history_changes niter. */
int hp_part_history(hp_part *part, double *hpropars, double *hcharges, double *hentropies, double *hchanges);

/* Per-atom inner-iteration counts and HP_ST_* status bits of the last outer
 * iteration, plus the OR of the status bits over all iterations. */
int hp_part_status(hp_part *part, int32_t *inner_niter, int32_t *status_last, int32_t *status_all);

/* Spherical averages of the AIM densities of the last update (gt1 only):
 * the "spherical_average_{a}" cache entries of mbis.py:190-198 (nshell). */
int hp_part_spherical_averages(hp_part *part, double *rhobar);

/* AIM weights of atom `atom` for the pro-atoms that produced the last update
 * ("at_weights_{a}" of stockholder.py:364-377): its own atom-grid points for
 * grid_type 1 (n = its atom-grid size), all points for grid_type 2 (n =
 * npoint). */
int hp_part_at_weights(hp_part *part, int atom, double *out, int64_t n);

/* Promolecule density of the last update ("promoldens", stockholder.py:355). */
int hp_part_promoldens(hp_part *part, double *out, int64_t n);

/* AIM multipole moments (core/base.py:329-402) of order lmax >= 0 in HORTON
 * order: cartesian natom*(lmax+1)(lmax+2)(lmax+3)/6, pure natom*(lmax+1)^2,
 * radial natom*(lmax+1) (pseudo_numbers added to the monopoles, electron sign
 * applied as the reference does).  Any of the three may be NULL.  Grid types
 * 1 and 2 integrate each atom's AIM density on its own atom grid with the
 * atom-grid weights; grid_type 3 on the whole molecular grid with the
 * molecular weights (local = False, core/base.py:455-459). */
int hp_part_moments(hp_part *part, int lmax, double *cartesian, double *pure, double *radial);

/* hp_part_moments without allocating: the same kernels, with their partial
 * sums in a static device workspace of 2^20 doubles (8 MiB) that each GPU
 * holds until the process ends.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under stream capture,
 * serving loops).  Molecules whose moments need more than the workspace get
 * HP_E_ARG: (partial rows + natom) x S sums plus the natom x Q outputs, with
 * one partial row per radial shell (grid_type 3: per atom and 4096-point
 * chunk), S = 40 for lmax <= 3 and S = Q otherwise, Q the number of
 * quantities of lmax (cartesian + pure + radial).  Every
 * call on one GPU uses that GPU's workspace: calls that overlap on a GPU
 * (several host threads) race, so serialise them. */
int hp_part_moments_static(hp_part *part, int lmax, double *cartesian, double *pure, double *radial);

/* ---- gLISA: the global problem (glisa.py:55-1028) ----
 * A partitioning created with HP_SCHEME_LISA and grid_type 2 or 3 holds the
 * basis functions g_i (fn_alpha, fn_order, fn_norm) at unit coefficient.
 * These entry points evaluate, on the molecular grid with the molecular
 * weights, what gLISA's solvers need per step; the solvers' own logic (the
 * sc fixed point, the Newton step and its line search) runs in the caller. */

/* At the coefficients c (sum K_a values): the promolecule rho0 = sum_i c_i g_i
 * (calc_promol_dens, glisa.py:346-348, no floors), the objective
 * f = sum_p w rho ln(rho/rho0) (also the entropy of stockholder.py:145-151 at
 * rho0), for nderiv >= 1 the gradient -sum_p w rho g_i / rho0 and for
 * nderiv = 2 the Hessian sum_p w rho g_i g_j / rho0^2 (M x M, row-major, on
 * the matrix cores) of _working_matrix (glisa.py:411-479), points where rho or
 * rho0 is below density_cutoff left out; min_promol receives min_p rho0.
 * Any output may be NULL. */
int hp_part_global_terms(hp_part *part, const double *coefs, int nderiv, double *objective, double *gradient,
                         double *hessian, double *min_promol);

/* Per atom over the molecular grid: the smallest value of its pro-atom at
 * coefs_new (is_proatom_valid, glisa.py:315-330) and
 * sum_p w (rho0_a(new) - rho0_a(old))^2 (compute_change on the molecular
 * grid, iterstock.py:32-45).  Either output may be NULL. */
int hp_part_global_proatoms(hp_part *part, const double *coefs_new, const double *coefs_old, double *min_proatom,
                            double *msd);

/* AIM populations sum_p w rho clip(rho0_a / promol) on the molecular grid,
 * from the last weights pass (hp_part_weights or hp_part_run): the charges
 * step of glisa.py:270-278.  grid_type 2 or 3.  The tabulated schemes give,
 * on grid types 1 and 2, the atom-grid populations of do_populations
 * (core/base.py:277-300): sum over atom a's own grid of the atom-grid weight
 * times w_a rho. */
int hp_part_populations(hp_part *part, double *populations);

/* Timing of one kernel (HP_KERNEL_*) measured with HIP events on the
 * library's stream: total milliseconds and launches, over the last
 * hp_part_run (HP_KERNEL_GLISA_HESS: since the partitioning was created). */
int hp_part_kernel_time(hp_part *part, int kernel, double *total_ms, int64_t *launches);

/* Algorithmic work of a partitioning, from its inputs (for rooflines). */
typedef struct {
  int64_t points;          /* grid points read by one weights pass */
  double evals_per_pass;   /* pro-atom function evaluations of one weights pass
                              (points x functions of the atoms that reach them) */
  double bytes_per_pass;   /* HBM bytes one weights pass must move: 48 per point
                              (x, y, z, density, angular and molecular weights)
                              + 16 per radial shell written */
  int64_t hessian_passes;  /* grid_type 2 Newton solvers, last hp_part_run: atom
                              passes over the grid that formed B diag(w) B^T */
  double hessian_flops;    /* ... their 2 K_a^2 N flops in total */
  double evals_done_fp32;  /* pro-atom function evaluations the weights passes of
                              the last hp_part_run performed in FP32 ... */
  double evals_done_fp64;  /* ... and in FP64 (after cut-off and tail skipping) */
} hp_part_work;

int hp_part_get_work(hp_part *part, hp_part_work *out);

#define HP_KERNEL_WEIGHTS 0 /* fused pro-atom/promolecule/weights/sphavg pass */
#define HP_KERNEL_INNER 1   /* per-atom inner solvers */
#define HP_KERNEL_LISA_GRID 2 /* molecular-grid Newton pass (grid_type 2) */
#define HP_KERNEL_GLISA_HESS 3 /* gLISA Hessian pass of hp_part_global_terms (accumulated since create) */

/* ---- continuous batching: many molecules in flight on one GPU ----
 *
 * The reference partitions one molecule per WPart object, one
 * do_partitioning() call after another (the part-dens driver loops over its
 * input files, scripts/program.py:120-131, one single_launch each,
 * scripts/partition_density.py:172-263).  A batch
 * keeps many such partitionings resident on the device and advances all of
 * them with the same launches: every tick is one outer iteration
 * (iterstock.py:159-193) of every molecule in flight.  Requests are admitted
 * while the GPU works, as slots and arena space free up; each one stops on
 * its own convergence test and leaves the batch.  The iteration of each
 * molecule is the same as hp_part_run's with the same precision (the
 * kernels see one molecule's data at a time, in the same order).
 *
 * Limits: grid_type 1, schemes MBIS (HP_SOLVER_MBIS) and GISA (HP_SOLVER_QP),
 * no cut-off radius.  The settings the kernels take per launch are shared by
 * every request (hp_batch_config): density_cutoff; inner_threshold and
 * inner_maxiter of MBIS requests; negative_cutoff and population_cutoff of
 * GISA requests.  Submissions with other values are rejected. */
typedef struct {
  int64_t point_capacity;    /* grid points resident at once */
  int32_t atom_capacity;     /* atoms resident at once (+1 per molecule) */
  int32_t shell_capacity;    /* radial shells resident at once (+1 per molecule) */
  int32_t function_capacity; /* pro-atom functions resident at once */
  int32_t max_inflight;      /* molecules resident at once */
  int64_t qp_capacity;       /* doubles of GISA QP matrices resident at once */
  int64_t max_molecule_points; /* largest molecule accepted (staging size) */
  int32_t ticks_per_step;    /* outer iterations enqueued per hp_batch_step */
  int32_t use_graphs;        /* replay the ticks of a step as a captured
                                hipGraph while the molecules in flight do
                                not change */
  double inner_threshold;    /* shared inner-solver settings, as in */
  int32_t inner_maxiter;     /* hp_scheme_desc */
  double density_cutoff;
  double negative_cutoff;
  double population_cutoff;
  int32_t groups;            /* molecule groups of equal weights cost, 1..4,
                                one stream each: the groups' weights launches
                                take turns, so that each group's per-atom
                                solves run beside the other groups' weights
                                passes.  0 selects 3 when the process has at
                                least 8 hardware queues (GPU_MAX_HW_QUEUES,
                                read when HIP starts), else 2: streams beyond
                                the queues share one and run in turn */
  int32_t inner_budget;      /* MBIS inner iterations per atom and launch
                                (0: no limit).  An atom whose inner solve needs
                                more keeps its state on the device and goes on
                                in the next ticks, while its molecule skips the
                                start of new outer iterations; the iterations
                                are the same, so results do not change. */
} hp_batch_config;

/* Why a request left the batch (hp_batch_result.event) */
#define HP_BATCH_CONVERGED 1 /* change < threshold */
#define HP_BATCH_MAXITER 2   /* iteration cap reached (iterstock.py:187) */
#define HP_BATCH_ERROR 3     /* a per-atom solver raised (status bits say which) */

typedef struct {
  int64_t ticket;  /* from hp_batch_submit */
  int32_t event;   /* HP_BATCH_* */
  int32_t niter;   /* outer iterations */
  int32_t status;  /* OR of the per-atom HP_ST_* bits over all iterations */
  int32_t status_last; /* fatal HP_ST_* bits of the last iteration */
  double change;   /* last change */
  double entropy;  /* last entropy */
} hp_batch_result;

typedef struct hp_batch hp_batch;

/* Allocate the device arena of a batch (capacities in cfg). */
int hp_batch_create(int device, const hp_batch_config *cfg, hp_batch **out);

/* Wait for the GPU, then release the batch and its device memory. */
int hp_batch_destroy(hp_batch *batch);

/* Queue one partitioning request.  mol/scheme as for hp_part_create,
 * propars0 the initial parameters (as hp_part_set_propars), precision and
 * n_polish as for hp_part_run.  The arrays behind mol, scheme and propars0
 * must stay valid and unchanged until the request's result is returned by
 * hp_batch_poll; propars_out (P doubles) and charges_out (natom) receive
 * the results before then.  *ticket identifies the request. */
int hp_batch_submit(hp_batch *batch, const hp_molecule_desc *mol, const hp_scheme_desc *scheme,
                    const double *propars0, int precision, int n_polish, double *propars_out, double *charges_out,
                    int64_t *ticket);

/* Advance the batch: wait for the previous step's iterations, retire the
 * requests that stopped, admit queued requests that fit, and enqueue the
 * next ticks_per_step iterations of everything in flight.  Returns the
 * number of requests resident on the device and still queued. */
int hp_batch_step(hp_batch *batch, int32_t *n_inflight, int32_t *n_queued);

/* Take up to max finished requests (oldest first); *n receives how many. */
int hp_batch_poll(hp_batch *batch, hp_batch_result *out, int32_t max, int32_t *n);

/* Work and GPU time of a batch since it was created (HIP events on the
 * batch's compute stream, read at each step for the ticks of the previous
 * one).  The counts are algorithmic: one molecule-iteration of a molecule
 * with N points and F pro-atom functions counts N points and N*F function
 * evaluations, whatever the kernels skip. */
typedef struct {
  int64_t ticks;           /* ticks run */
  int64_t iterations;      /* molecule-iterations done */
  int64_t completed;       /* requests finished */
  double tick_ms;          /* GPU time of the ticks */
  double weights_ms;       /* ... of their weights launches (pro-atoms, promolecule,
                              weights); 0 with use_graphs, which times whole steps */
  double inner_ms;         /* ... of their per-atom solver launches (0 with use_graphs) */
  double points_fp32;      /* grid points through the FP32 weights pass */
  double points_fp64;      /* grid points through the FP64 weights pass */
  double evals_fp32;       /* pro-atom function evaluations in FP32 */
  double evals_fp64;       /* pro-atom function evaluations in FP64 */
  double done_fp32;        /* function evaluations the FP32 weights launches
                              performed (after tail skipping) ... */
  double done_fp64;        /* ... and the FP64 ones; hp_batch_get_stats waits
                              for the work enqueued so far to read them */
} hp_batch_stats;

int hp_batch_get_stats(hp_batch *batch, hp_batch_stats *out);

/* Page-locked host memory for request arrays: the batch's uploads from it
 * run asynchronously, overlapping the ticks (no reference counterpart; a
 * serving front end keeps its incoming grids and densities there). */
int hp_host_alloc(int64_t bytes, void **ptr);
int hp_host_free(void *ptr);

/* ---- grid utilities (input preparation, not the iteration) ---- */

/* Becke fuzzy-cell weights (qc-grid BeckeWeights, used by MolGrid in
 * tests/test_wpart.py:50-51): for every point, the weight of the atom whose
 * atom grid holds it (atom_offsets: natom+1 point offsets).  radii in bohr. */
int hp_becke_weights(int device, int natom, const double *coordinates, const double *radii, int64_t npoint,
                     const double *points, const int64_t *atom_offsets, double *out);

/* sum_a sum_k coef_ak exp(-alpha_ak |p - R_a|) at every point: the
 * Slater-shell densities of the synthetic workloads (no reference
 * counterpart; the reference reads densities from wavefunction files,
 * scripts/generate_density.py:119-204).  With non-negative coefficients and
 * exponents, atoms whose terms are below 2^-64 of the density at every point
 * of a block of 256 consecutive points are skipped. */
int hp_slater_density(int device, int natom, const double *coordinates, const int32_t *fn_offsets,
                      const double *coef, const double *alpha, int64_t npoint, const double *points, double *out);

/* ---- measurement probes (no reference counterpart) ---- */

/* Device-to-device copy of nbytes, reps times: achieved GB/s (read + write
 * bytes per second) and milliseconds per copy.  The HBM bandwidth that a
 * streaming kernel reaches on this device, and the reference point for
 * calibrating memory counters. */
int hp_stream_copy(int device, int64_t nbytes, int reps, double *gbps, double *ms_per_copy);

/* Exponentials per second of the weights pass's exp (precision
 * HP_PREC_FP64: the FP64 exp; HP_PREC_MIXED: the FP32 exp), independent
 * arguments in [-30, 0] on every lane of the device. */
int hp_exp_rate(int device, int precision, double *exps_per_s);

#ifdef __cplusplus
}
#endif

#endif /* HORTONPART_H */
