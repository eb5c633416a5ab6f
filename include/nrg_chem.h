/*
 * nrg_chem.h -- C-ABI of the MI355X (gfx950) chemistry step for NRG.
 *
 * This library replaces the body of NRG's `chemical_kinetics_solver` type
 * (computing_module/src/current_build/chemical_kinetics_solver.f90:81-103):
 * the per-cell operator-split stiff chemistry step that the CABARET, FDS
 * low-Mach and CPM solvers call once per flow time step.  A Fortran
 * ISO_C_BINDING module (nrg_amd/fortran/nrg_chem_binding.f90) binds these
 * symbols, so the callers keep calling
 *     chem_kin_solver = chemical_kinetics_solver_c(manager)
 *     call chem_kin_solver%solve_chemical_kinetics(time_step)
 * unchanged.
 *
 * Conventions
 *  - Plain C types only; every entry point returns an `int` status
 *    (NRG_CHEM_OK == 0) and never aborts the process.  nrg_chem_strerror()
 *    turns a status into text; nrg_chem_last_error() gives the detail of the
 *    last failure on the calling thread.
 *  - Arrays follow NRG's Fortran layout: column-major, 1-based indices in the
 *    index tables, fields allocated with ghost layers (see nrg_grid).
 *  - Entry points ending in `_dev` take device (HBM) pointers and run on the
 *    handle's HIP stream; the others take host pointers and stage through
 *    device buffers owned by the handle.  All entry points synchronise their
 *    stream before returning.
 *  - A handle is not re-entrant: one host thread calls it at a time, as the
 *    reference calls solve_chemical_kinetics from the master thread only
 *    (computing_module/src/current_build/cabaret_solver.f90:1362).
 */
#ifndef NRG_CHEM_H
#define NRG_CHEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NRG_CHEM_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum {
  NRG_CHEM_OK = 0,
  NRG_CHEM_E_ARG = 1,          /* bad argument (null pointer, bad size, bad bounds)          */
  NRG_CHEM_E_MECH = 2,         /* mechanism not supported by the compiled kernels             */
  NRG_CHEM_E_HIP = 3,          /* HIP runtime error (allocation, launch, copy)                */
  NRG_CHEM_E_NO_DEVICE = 4,    /* no gfx950 device available                                  */
  NRG_CHEM_E_FORMAT = 5,       /* a mechanism / thermo / molar-mass file is unreadable        */
  NRG_CHEM_E_STEP_ZERO = 41,   /* a cell's step size went to zero  (SLATEC IERFLG 41)         */
  NRG_CHEM_E_REPEATED_FAIL = 42/* a cell failed MXTRY=50 attempts (SLATEC IERFLG 42)          */
};

/* ---- per-cell status written to `cell_status` (nullable) ---------------- */
enum {
  NRG_CELL_OK = 0,             /* integrated to dt                                            */
  NRG_CELL_GATED = 1,          /* skipped: T < t_gate or bc_marker != 0 (outputs zeroed)      */
  NRG_CELL_MXSTEP = 2,         /* MXSTEP steps taken before dt: partial integration, as the
                                  reference returns it silently (slatec/ddriv3.f90:1329-1341)  */
  NRG_CELL_STEP_ZERO = 3,      /* step size underflow; the reference STOPs here (IERFLG 41)   */
  NRG_CELL_REPEATED_FAIL = 4,  /* 50 failed attempts; the reference STOPs here (IERFLG 42)    */
  NRG_CELL_SINGULAR = 5        /* exactly singular iteration matrix (unreachable in practice)  */
};

/* Jacobian used by the BDF corrector. */
enum {
  NRG_JAC_FD_SLATEC = 0,       /* SLATEC DDPST MITER=2 finite differences: reference parity   */
  NRG_JAC_ANALYTIC = 1         /* analytic d(omega)/dC at fixed k: fast mode                  */
};

/*
 * Queue order of the cells (same arithmetic, same results).  STIFFNESS bins
 * cells by the step count each took in the handle's previous call on the same
 * grid (by temperature on the first call) and hands out the most expensive
 * bins first: a wave then holds cells of similar cost, and the long cells
 * start early instead of forming the tail.
 */
enum {
  NRG_ORDER_NATURAL = 0,       /* inner-loop order                                            */
  NRG_ORDER_STIFFNESS = 1
};

/* Launch form of the per-cell integrator (same arithmetic, same results). */
enum {
  NRG_FORM_PLAIN = 0,          /* one cell per lane, one launch over all inner cells          */
  NRG_FORM_PERSISTENT = 1      /* resident grid; lanes refill from a work queue as cells end  */
};

/*
 * Mechanism and thermochemistry, exactly the arrays NRG's
 * chemical_properties (package_library/src/chemical_properties_class.f90:15-52)
 * and thermophysical_properties (package_library/src/thermophysical_properties_class.f90:155-169)
 * hold after set_properties, in Fortran order:
 *   reactions_type(R)               int32
 *   chemical_coeffs(4,R,2)          int32: (1,r,side) = number of terms, (2..4,r,side) = 1-based
 *                                   species index, S+1 = third body M, 0 = unused; side 1 = left
 *   A, beta, E_act, A_low, beta_low, E_act_low (R)   already unit-scaled (:162-180)
 *   Troe_coeffs(R,4)                (alpha, T***, T*, T**)
 *   enhanced_efficiencies(R,S)
 *   a_coeffs(7,2,S)                 NASA-7, multiplied by R_u (:167); set 1 = T < 1000 K
 *   molar_masses(S)                 kg/mol
 */
typedef struct nrg_chem_mech {
  int32_t species_number;
  int32_t reactions_number;
  const int32_t *reactions_type;
  const int32_t *chemical_coeffs;
  const double *A, *beta, *E_act;
  const double *A_low, *beta_low, *E_act_low;
  const double *Troe_coeffs;
  const double *enhanced_efficiencies;
  const double *a_coeffs;
  const double *molar_masses;
} nrg_chem_mech;

/*
 * Integrator options.  nrg_chem_default_opts() fills the reference literals of
 * computing_module/src/current_build/chemical_kinetics_solver.f90:270-290:
 * EPS = 1e-9, EWT = 1e-4 (IERROR = 3), HMAX = 1.0E-7 (a default-real literal,
 * i.e. (double)1.0e-7f), MXSTEP = 10000, MXORD = 5, and the T >= 305 K gate of :228.
 */
typedef struct nrg_chem_opts {
  double eps;
  double ewt;
  double hmax;
  double t_gate;
  int32_t mxstep;
  int32_t mxord;
  int32_t jacobian;            /* NRG_JAC_*                                                    */
  int32_t device;              /* HIP device ordinal                                           */
  int32_t form;                /* NRG_FORM_*                                                   */
  int32_t jac_batch_pct;       /* a wave runs a batched Jacobian when at least this percentage
                                  of its pending lanes wait for one (1..100; 0 = default 50)   */
  int32_t waves_per_cu;        /* persistent form: resident waves per CU (0 = occupancy limit) */
  int32_t cell_order;          /* NRG_ORDER_*: order in which the queue hands out cells        */
  int32_t use_graph;           /* 1: capture the step's launches in a HIP graph and replay it
                                  while the arguments stay the same (same results)             */
  int32_t chunk_log2;          /* queue chunk = 2^chunk_log2 cells (16..26; 0 = 23): cells per
                                  integrator launch and per cell-record buffer (496 B per cell
                                  for KEROMNES: 4.2 GB at 2^23)                                */
  int32_t errtest_gather_pct;  /* 0..100: a lane whose corrector has converged waits at the
                                  error test until this percentage of the wave's lanes that
                                  are stepping is there (100: the wave's cells take their
                                  steps' error tests together; 0: no waiting).  Same results. */
  int32_t refill_min_idle;     /* persistent form: a wave hands new cells to its idle lanes once
                                  this many of its 64 lanes are idle, or all are (1..64; 0 =
                                  default 63: the wave's cells start together).  Same results. */
} nrg_chem_opts;

/*
 * Field geometry.  Every field is a column-major FP64 (or int8 for markers)
 * array with Fortran bounds lo(d):hi(d); the step visits the inner cells
 * inner_lo(d):inner_hi(d) (computational_domain%get_local_inner_cells_bounds,
 * package_library/src/computational_domain_class.f90:488-503).  Cell (i,j,k)
 * lives at offset (i-lo0) + (hi0-lo0+1)*((j-lo1) + (hi1-lo1+1)*(k-lo2)).
 */
typedef struct nrg_grid {
  int32_t lo[3], hi[3];
  int32_t inner_lo[3], inner_hi[3];
} nrg_grid;

/* Work counters summed over the integrated cells of one call (SLATEC semantics:
 * NFE counts the N extra RHS calls of each finite-difference Jacobian,
 * slatec/ddpst.f90:110). */
typedef struct nrg_chem_stats {
  int64_t cells_inner;
  int64_t cells_integrated;
  int64_t nstep_sum, nfe_sum, nje_sum, ncorr_sum, natt_sum;
  int32_t nstep_max, nfe_max;
  int64_t status_count[8];     /* indexed by NRG_CELL_*                                        */
  double kernel_ms;            /* device time of the chemistry kernel(s), HIP events           */
} nrg_chem_stats;

typedef struct nrg_chem_handle nrg_chem_t;

void nrg_chem_default_opts(nrg_chem_opts *opts);
const char *nrg_chem_strerror(int status);
const char *nrg_chem_last_error(void);
int nrg_chem_abi_version(void);

/*
 * Constructor.  Replaces chemical_kinetics_solver_c(manager)
 * (chemical_kinetics_solver.f90:111-173): copies the mechanism coefficients to
 * the device, precomputes the Gear BDF tables (slatec/ddcst.f90) and allocates
 * nothing per cell.  `opts` may be NULL (defaults).
 */
int nrg_chem_create(const nrg_chem_mech *mech, const nrg_chem_opts *opts, nrg_chem_t **out);
void nrg_chem_destroy(nrg_chem_t *h);

/*
 * One chemistry step on device-resident fields.  Replaces
 * solve_chemical_kinetics(this, time_step) (chemical_kinetics_solver.f90:175-358):
 * for every inner cell with T >= t_gate and bc == 0, integrates
 * d[C]/dt = omega([C]) at the cell's fixed T from 0 to dt with the reference's
 * Gear-BDF settings, and writes
 *   Y_prod[s]  = ((C_out - C_in) * W_s) / dt      [kg m^-3 s^-1]
 *   Y_prod2[s] =  C_out - C_in                     [mol m^-3]
 *   E_f_prod   = -sum_s (h_s(T_ref) * (C_out - C_in)) / dt   [J m^-3 s^-1]
 * Gated cells get Y_prod = 0, E_f_prod = 0 and keep Y_prod2 unchanged, as the
 * reference does (:341-346).  `p` (pressure) is accepted for interface
 * fidelity and is unused, as in the reference.  Y, Y_prod, Y_prod2 are arrays
 * of S field pointers (one allocation per species, Y%pr(s)%cells).
 * `bc` and `cell_status` and `p` may be NULL (no markers / no status).
 * Returns NRG_CHEM_E_STEP_ZERO / _REPEATED_FAIL when a cell hit a condition
 * that makes the reference STOP; the other cells' outputs are still written.
 */
int nrg_chem_step_dev(nrg_chem_t *h, const nrg_grid *grid, double dt,
                      const double *T, const double *p, const double *rho,
                      const double *const *Y, const int8_t *bc,
                      double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                      int32_t *cell_status, nrg_chem_stats *stats);

/*
 * nrg_chem_step_dev plus per-cell work counters, the per-call IWORK statistics
 * of DDRIV3 (slatec/ddriv3.f90 IWORK(3..5): NSTEP, NFE, NJE), one int32 per
 * inner cell in inner-loop order (i fastest).  All three or none may be NULL.
 */
int nrg_chem_step_dev_diag(nrg_chem_t *h, const nrg_grid *grid, double dt,
                           const double *T, const double *p, const double *rho,
                           const double *const *Y, const int8_t *bc,
                           double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                           int32_t *cell_status, nrg_chem_stats *stats,
                           int32_t *nstep, int32_t *nfe, int32_t *nje);

/*
 * nrg_chem_step_dev that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot call hipMalloc (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  Same arguments,
 * kernels, arithmetic and outputs as nrg_chem_step_dev.  The work counters,
 * queue heads and the rate constants of one chunk live in a fixed device
 * array of NRG_NOALLOC_SCRATCH_BYTES that the library holds on every device
 * it runs on until the process ends; the grid is integrated in chunks of
 * NRG_NOALLOC_CHUNK_CELLS cells in natural cell order (the handle's cell_order
 * and use_graph are not applied; results are the same).  A grid of more than
 * NRG_NOALLOC_MAX_CELLS inner cells returns NRG_CHEM_E_ARG before any work.
 * All handles on one device share that array and each call leaves its
 * contents behind: calls that overlap on one device (other host threads,
 * other handles' streams) race and must be serialised by the caller.
 */
#define NRG_NOALLOC_SCRATCH_BYTES (64u << 20)          /* 64 MiB per device             */
#define NRG_NOALLOC_CHUNK_CELLS (1 << 17)              /* cells per integrator launch   */
#define NRG_NOALLOC_MAX_CELLS (48 * NRG_NOALLOC_CHUNK_CELLS)  /* 6,291,456 inner cells  */
int nrg_chem_step_dev_noalloc(nrg_chem_t *h, const nrg_grid *grid, double dt,
                              const double *T, const double *p, const double *rho,
                              const double *const *Y, const int8_t *bc,
                              double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                              int32_t *cell_status, nrg_chem_stats *stats);

/* Same contract with host pointers (fields are staged through device buffers). */
int nrg_chem_step(nrg_chem_t *h, const nrg_grid *grid, double dt,
                  const double *T, const double *p, const double *rho,
                  const double *const *Y, const int8_t *bc,
                  double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                  int32_t *cell_status, nrg_chem_stats *stats);

/*
 * Split-step closure on device: the legacy CABARET source application
 * (computing_module/src/current_build/cabaret_solver.f90:3316-3318, 3386-3416)
 * followed by the e->T inversion of apply_state_equation
 * (computing_module/src/current_build/table_approximated_real_gas.f90:410-440,
 * package_library/src/thermophysical_properties_class.f90:409-452), for cells with bc == 0:
 *   E_f += (E_f_prod * dt) / rho ;  Y_s += (Y_prod_s * dt) / rho ;
 *   Y_s  = max(Y_s, 0) / sum_s(Y_s before clipping)
 *   e    = E_f - 0.5*|v|^2 ; W = 1/sum(Y/W_s) ; X_s = (Y_s/W_s)*W ; e_m = e*W
 *   T    = frozen-cp Newton on e_m from the old T ;  p = ((T*rho)*R_u)/W
 * `v` is an array of `ndim` velocity-component pointers (NULL/0: at rest);
 * `p` may be NULL.  E_f, Y, T, p are updated in place.
 */
int nrg_chem_closure_dev(nrg_chem_t *h, const nrg_grid *grid, double dt,
                         const double *rho, const double *const *v, int32_t ndim,
                         const double *E_f_prod, const double *const *Y_prod,
                         const int8_t *bc, double *E_f, double *const *Y,
                         double *T, double *p);

/* e->T only (apply_state_equation without the source application). */
int nrg_state_equation_dev(nrg_chem_t *h, const nrg_grid *grid,
                           const double *rho, const double *const *v, int32_t ndim,
                           const double *E_f, const double *const *Y, const int8_t *bc,
                           double *T, double *p);

/*
 * Mixture sensible+kinetic energy per unit mass consistent with T and Y, the
 * inverse of the e->T step: E_f = (sum_s X_s*(e_s(T) - h_s(T_ref)))/W + 0.5*|v|^2.
 * Used to set up states; the reference builds E_f the same way from T
 * (thermophysical_properties_class.f90:288-359).
 */
int nrg_energy_from_temperature_dev(nrg_chem_t *h, const nrg_grid *grid,
                                    const double *T, const double *const *Y,
                                    const double *const *v, int32_t ndim, double *E_f);

/*
 * Autotuner of the batched form.  A cell is one lane and a workgroup one
 * wave (64 cells), so what is tuned is how the queue feeds the waves: launch
 * form, cell order, the error-test gathering, the batched-Jacobian threshold,
 * the refill threshold, resident waves per CU and the queue chunk size.  Starting from the handle's current set-up it tries
 * one change at a time (greedy), each as a full chemistry step on the given
 * fields timed with HIP events, keeps the fastest and leaves it in the
 * handle.  Every trial writes the step's outputs (identical for every set-up).
 * Arguments as nrg_chem_step_dev; `out` receives the trials and the choice.
 */
#define NRG_TUNE_MAX_TRIALS 24
typedef struct nrg_chem_tune_trial {
  int32_t form, cell_order, jac_batch_pct, waves_per_cu;
  int64_t chunk_cells;
  double kernel_ms;
  int32_t errtest_gather_pct, refill_min_idle;
} nrg_chem_tune_trial;

typedef struct nrg_chem_tuning {
  nrg_chem_tune_trial best;
  int32_t n_trials;
  int32_t reserved;
  nrg_chem_tune_trial trials[NRG_TUNE_MAX_TRIALS];
} nrg_chem_tuning;

int nrg_chem_autotune_dev(nrg_chem_t *h, const nrg_grid *grid, double dt,
                          const double *T, const double *p, const double *rho,
                          const double *const *Y, const int8_t *bc,
                          double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                          nrg_chem_tuning *out);

/*
 * Table-approximated chemistry: the reference's ODE_solver = 'table_approximated'
 * mode (computing_module/src/current_build/chemical_kinetics_solver.f90:168-169,
 * 250-346, interpolate_chemical_kinetics :650-685).
 *
 * nrg_chem_set_table replaces read_chemical_kinetics_table (:608-648): `rows`
 * (2 .. 2^24)
 * temperatures T_t and, species-major, dC_t[s*rows + i] (the table's columns
 * 2..S+1, i.e. Y_prod2 of the run that wrote it, write_chemical_kinetics_table
 * :542-606).  The table is copied to the device.
 *
 * nrg_chem_table_step_dev has nrg_chem_step_dev's arguments and gate (p and
 * rho are not read).  For an integrated cell: C_in = Y; C_out = the table
 * row pair bracketing T, interpolated linearly (the last bracketing interval
 * wins; T below the table gives 0, above it the last row; no bracketing
 * interval leaves C_out = Y); Y_prod = C_out*W/dt, Y_prod2 = C_out - C_in,
 * E_f_prod = -sum_s h_s(T_ref)*C_out/dt, status NRG_CELL_OK.  stats carries
 * cells_inner, cells_integrated and kernel_ms.
 */
int nrg_chem_set_table(nrg_chem_t *h, int32_t rows, const double *T_t, const double *dC_t);
/* Host-array form of nrg_chem_table_step_dev for callers whose fields live in
 * host memory (the reference's own data manager): staged like nrg_chem_step. */
int nrg_chem_table_step(nrg_chem_t *h, const nrg_grid *grid, double dt,
                        const double *T, const double *p, const double *rho,
                        const double *const *Y, const int8_t *bc,
                        double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                        int32_t *cell_status, nrg_chem_stats *stats);
int nrg_chem_table_step_dev(nrg_chem_t *h, const nrg_grid *grid, double dt,
                            const double *T, const double *p, const double *rho,
                            const double *const *Y, const int8_t *bc,
                            double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                            int32_t *cell_status, nrg_chem_stats *stats);

/* HIP stream the handle launches on (a hipStream_t), for callers that time or order work. */
void *nrg_chem_stream(nrg_chem_t *h);

/*
 * Name of the reaction topology whose kernels serve the handle ("KEROMNES",
 * "ZHANG", "TEREZA", "KONNOV", "ACETYLENE"), or NULL for a NULL handle.  No counterpart in the reference, which
 * interprets any mechanism at run time: here the kernels unroll a mechanism's
 * reaction lists at compile time, one build per topology, and nrg_chem_create
 * picks the build whose reaction types and species lists
 * (chemical_properties_class.f90:15-52) equal the mechanism's.
 */
const char *nrg_chem_topology(nrg_chem_t *h);

#ifdef __cplusplus
}
#endif
#endif /* NRG_CHEM_H */
